// Scene files on the host side (include/pyre_gen.h): the reference CLI's
// --scene (.pbrt, .gltf/.glb) and --env (.hdr) inputs turned into the
// pyre_preset container that pyre_scene_upload consumes
// (crates/pyre-cli/src/main.rs:114-166).
//
// Restated from the reference:
//   PBRT     crates/pyre/src/io/pbrt/lexer.rs, parser.rs, builder.rs, mod.rs
//   glTF     crates/pyre/src/io/gltf.rs (over the gltf 1.4.1 crate's import)
//   Radiance crates/pyre/src/io/hdr.rs (over the image 0.25.9 crate's decoder)
// Transforms follow glam 0.29.3's f32 Mat4 / Mat3 / Quat arithmetic operation
// by operation (column-major, left-to-right sums, no fused multiply-add:
// this file is built with -ffp-contract=off), so baked vertices are the
// reference's bits.
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <utility>
#include <vector>

#include "pmath.hpp"
#include "preset.hpp"

using namespace pyre;

namespace {

// ---------------------------------------------------------------- errors --

struct LoadError {
  int code;
  std::string msg;
};

[[noreturn]] void fail(int code, std::string msg) { throw LoadError{code, std::move(msg)}; }

// Rust's {:?} of a string: quoted, with \" \\ \n \r \t escaped.
std::string dbg(const std::string& s) {
  std::string o = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') {
      o += '\\';
      o += c;
    } else if (c == '\n') {
      o += "\\n";
    } else if (c == '\r') {
      o += "\\r";
    } else if (c == '\t') {
      o += "\\t";
    } else {
      o += c;
    }
  }
  return o + "\"";
}

std::vector<uint8_t> read_bytes(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) fail(PYRE_EIO, "io: " + std::string(std::strerror(errno)) + " (" + path + ")");
  std::vector<uint8_t> out;
  uint8_t buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.insert(out.end(), buf, buf + n);
  const bool bad = std::ferror(f) != 0;
  std::fclose(f);
  if (bad) fail(PYRE_EIO, "io: read error (" + path + ")");
  return out;
}

std::string dir_of(const std::string& path) {
  const size_t s = path.find_last_of('/');
  return s == std::string::npos ? std::string(".") : path.substr(0, s == 0 ? 1 : s);
}

// Path::join: an absolute second part replaces the first.
std::string join(const std::string& base, const std::string& rel) {
  if (!rel.empty() && rel[0] == '/') return rel;
  if (base.empty() || base.back() == '/') return base + rel;
  return base + "/" + rel;
}

std::string lower_ext(const std::string& path) {
  const size_t slash = path.find_last_of('/');
  const size_t dot = path.find_last_of('.');
  if (dot == std::string::npos || (slash != std::string::npos && dot < slash) || dot == path.size() - 1) return "";
  std::string e = path.substr(dot + 1);
  for (char& c : e) c = (char)std::tolower((unsigned char)c);
  return e;
}

// ------------------------------------------------------- glam arithmetic --

struct M4 {
  float c[4][4];  // c[column][row]
};

M4 m4_identity() {
  M4 m{};
  for (int i = 0; i < 4; ++i) m.c[i][i] = 1.0f;
  return m;
}

// Mat4 * Vec4 (glam mul_vec4): ((x * v.x + y * v.y) + z * v.z) + w * v.w.
void m4_vec(const M4& a, const float v[4], float out[4]) {
  for (int r = 0; r < 4; ++r) {
    float s = a.c[0][r] * v[0];
    s = s + a.c[1][r] * v[1];
    s = s + a.c[2][r] * v[2];
    s = s + a.c[3][r] * v[3];
    out[r] = s;
  }
}

M4 m4_mul(const M4& a, const M4& b) {
  M4 m;
  for (int j = 0; j < 4; ++j) m4_vec(a, b.c[j], m.c[j]);
  return m;
}

M4 m4_translation(float x, float y, float z) {
  M4 m = m4_identity();
  m.c[3][0] = x;
  m.c[3][1] = y;
  m.c[3][2] = z;
  return m;
}

M4 m4_scale(float x, float y, float z) {
  M4 m{};
  m.c[0][0] = x;
  m.c[1][1] = y;
  m.c[2][2] = z;
  m.c[3][3] = 1.0f;
  return m;
}

// Mat4::from_quat (glam quat_to_axes).
M4 m4_from_quat(float x, float y, float z, float w) {
  const float x2 = x + x, y2 = y + y, z2 = z + z;
  const float xx = x * x2, xy = x * y2, xz = x * z2;
  const float yy = y * y2, yz = y * z2, zz = z * z2;
  const float wx = w * x2, wy = w * y2, wz = w * z2;
  M4 m{};
  m.c[0][0] = 1.0f - (yy + zz);
  m.c[0][1] = xy + wz;
  m.c[0][2] = xz - wy;
  m.c[1][0] = xy - wz;
  m.c[1][1] = 1.0f - (xx + zz);
  m.c[1][2] = yz + wx;
  m.c[2][0] = xz + wy;
  m.c[2][1] = yz - wx;
  m.c[2][2] = 1.0f - (xx + yy);
  m.c[3][3] = 1.0f;
  return m;
}

// Mat4::transform_point3: x * p.x, then y * p.y + that, z * p.z + that, w + that.
V3 m4_point(const M4& m, V3 p) {
  float r[3];
  for (int i = 0; i < 3; ++i) {
    float s = m.c[0][i] * p.x;
    s = m.c[1][i] * p.y + s;
    s = m.c[2][i] * p.z + s;
    s = m.c[3][i] + s;
    r[i] = s;
  }
  return v3(r[0], r[1], r[2]);
}

// Mat4::transform_vector3.
V3 m4_vector(const M4& m, V3 v) {
  float r[3];
  for (int i = 0; i < 3; ++i) {
    float s = m.c[0][i] * v.x;
    s = m.c[1][i] * v.y + s;
    s = m.c[2][i] * v.z + s;
    r[i] = s;
  }
  return v3(r[0], r[1], r[2]);
}

// ----------------------------------------------------------- PBRT lexer --
// lexer.rs:54-180

struct Pos {
  uint32_t line, col;
};

std::string pos_str(Pos p) { return "line " + std::to_string(p.line) + ", col " + std::to_string(p.col); }

enum TokKind { T_IDENT, T_STRING, T_INT, T_FLOAT, T_LBRACKET, T_RBRACKET };

struct Token {
  TokKind kind;
  std::string s;
  int64_t i = 0;
  double f = 0.0;
  Pos pos;
};

bool is_number_start(uint8_t c) { return (c >= '0' && c <= '9') || c == '-' || c == '+' || c == '.'; }
bool is_number_body(uint8_t c) { return is_number_start(c) || c == 'e' || c == 'E'; }
bool is_alpha(uint8_t c) { return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z'); }
bool is_ident_start(uint8_t c) { return is_alpha(c) || c == '_'; }
bool is_ident_body(uint8_t c) { return is_alpha(c) || (c >= '0' && c <= '9') || c == '_'; }

// Rust's i64::from_str: optional sign, then at least one decimal digit.
bool parse_i64(const std::string& t, int64_t* v) {
  size_t k = (t[0] == '+' || t[0] == '-') ? 1 : 0;
  if (k == t.size()) return false;
  for (size_t j = k; j < t.size(); ++j)
    if (t[j] < '0' || t[j] > '9') return false;
  errno = 0;
  char* end = nullptr;
  const long long r = std::strtoll(t.c_str(), &end, 10);
  if (errno == ERANGE || *end) return false;
  *v = (int64_t)r;
  return true;
}

// Rust's f64::from_str over the lexer's number characters: correctly
// rounded, the whole text or nothing (overflow gives inf, underflow 0).
bool parse_f64(const std::string& t, double* v) {
  size_t k = (t[0] == '+' || t[0] == '-') ? 1 : 0;
  bool digits = false;
  while (k < t.size() && t[k] >= '0' && t[k] <= '9') digits = true, ++k;
  if (k < t.size() && t[k] == '.') {
    ++k;
    while (k < t.size() && t[k] >= '0' && t[k] <= '9') digits = true, ++k;
  }
  if (!digits) return false;
  if (k < t.size() && (t[k] == 'e' || t[k] == 'E')) {
    ++k;
    if (k < t.size() && (t[k] == '+' || t[k] == '-')) ++k;
    bool ed = false;
    while (k < t.size() && t[k] >= '0' && t[k] <= '9') ed = true, ++k;
    if (!ed) return false;
  }
  if (k != t.size()) return false;
  *v = std::strtod(t.c_str(), nullptr);
  return true;
}

bool valid_utf8(const uint8_t* p, size_t n) {
  size_t i = 0;
  while (i < n) {
    const uint8_t c = p[i];
    int len = c < 0x80 ? 1 : (c >> 5) == 6 ? 2 : (c >> 4) == 14 ? 3 : (c >> 3) == 30 ? 4 : 0;
    if (!len || i + len > n) return false;
    for (int k = 1; k < len; ++k)
      if ((p[i + k] & 0xC0) != 0x80) return false;
    i += len;
  }
  return true;
}

std::vector<Token> tokenize(const std::vector<uint8_t>& in) {
  std::vector<Token> out;
  const size_t n = in.size();
  size_t i = 0, line_start = 0;
  uint32_t line = 1;
  auto lex_err = [](Pos p, const std::string& what) { fail(PYRE_EINVAL, "lex: " + pos_str(p) + ": " + what); };
  while (i < n) {
    const Pos pos{line, (uint32_t)(i - line_start) + 1};
    const uint8_t c = in[i];
    if (c == ' ' || c == '\t' || c == '\r') {
      ++i;
    } else if (c == '\n') {
      ++line;
      ++i;
      line_start = i;
    } else if (c == '#') {
      while (i < n && in[i] != '\n') ++i;
    } else if (c == '[' || c == ']') {
      Token t{c == '[' ? T_LBRACKET : T_RBRACKET, "", 0, 0.0, pos};
      out.push_back(t);
      ++i;
    } else if (c == '"') {
      ++i;
      const size_t start = i;
      while (i < n && in[i] != '"') {
        if (in[i] == '\n') lex_err(pos, "unterminated string literal");
        ++i;
      }
      if (i >= n) lex_err(pos, "unterminated string literal");
      if (!valid_utf8(in.data() + start, i - start)) lex_err(pos, "string is not UTF-8");
      out.push_back(Token{T_STRING, std::string((const char*)in.data() + start, i - start), 0, 0.0, pos});
      ++i;
    } else if (is_number_start(c)) {
      const size_t start = i;
      while (i < n && is_number_body(in[i])) ++i;
      const std::string text((const char*)in.data() + start, i - start);
      Token t{T_FLOAT, "", 0, 0.0, pos};
      const bool floaty = text.find_first_of(".eE") != std::string::npos;
      if (!floaty && parse_i64(text, &t.i)) {
        t.kind = T_INT;
      } else if (!parse_f64(text, &t.f)) {
        lex_err(pos, "invalid number " + dbg(text));
      }
      out.push_back(t);
    } else if (is_ident_start(c)) {
      const size_t start = i;
      while (i < n && is_ident_body(in[i])) ++i;
      out.push_back(Token{T_IDENT, std::string((const char*)in.data() + start, i - start), 0, 0.0, pos});
    } else {
      std::string ch;
      if (c < 0x80) {
        ch = std::string("'") + (c == '\'' || c == '\\' ? "\\" : "") + (char)c + "'";
      } else {
        char b[16];
        std::snprintf(b, sizeof b, "'\\x%02x'", c);
        ch = b;
      }
      lex_err(pos, "unexpected character " + ch);
    }
  }
  return out;
}

// ---------------------------------------------------------- PBRT parser --
// parser.rs:14-357

enum PosKind { P_FLOAT, P_STRING, P_ARRAY };

struct Positional {
  PosKind kind;
  double f = 0.0;
  std::string s;
  std::vector<double> arr;
};

enum ValKind { V_FLOATS, V_INTS, V_STRINGS };

struct Param {
  std::string type, name;
  ValKind kind;
  std::vector<double> floats;
  std::vector<int64_t> ints;
  std::vector<std::string> strings;
  Pos pos;
};

struct Directive {
  std::string keyword;
  std::vector<Positional> positional;
  std::vector<Param> params;
  Pos pos;
};

const char* tok_kind(TokKind k) {
  switch (k) {
    case T_IDENT: return "identifier";
    case T_STRING: return "string";
    case T_INT: return "integer";
    case T_FLOAT: return "float";
    case T_LBRACKET: return "`[`";
    default: return "`]`";
  }
}

bool is_float_type(const std::string& t) {
  static const char* k[] = {"float",   "rgb",     "color",  "spectrum", "blackbody", "point",   "point2",
                            "point3",  "vector",  "vector2", "vector3", "normal",    "normal3", "xyz"};
  for (const char* s : k)
    if (t == s) return true;
  return false;
}

[[noreturn]] void parse_err(Pos p, const std::string& what) { fail(PYRE_EINVAL, "parse: " + pos_str(p) + ": " + what); }

[[noreturn]] void expected(Pos p, const std::string& what, const std::string& found) {
  parse_err(p, "expected " + what + ", found " + found);
}

bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\f' || c == '\v'; }

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && is_ws(s[a])) ++a;
  while (b > a && is_ws(s[b - 1])) --b;
  return s.substr(a, b - a);
}

std::vector<Directive> parse(const std::vector<Token>& toks) {
  std::vector<Directive> out;
  size_t i = 0;
  const Pos last = toks.empty() ? Pos{1, 1} : toks.back().pos;
  while (i < toks.size()) {
    const Token& head = toks[i++];
    if (head.kind != T_IDENT) expected(head.pos, "directive keyword", tok_kind(head.kind));
    Directive d;
    d.keyword = head.s;
    d.pos = head.pos;
    while (i < toks.size()) {
      const Token& t = toks[i];
      if (t.kind == T_IDENT) break;
      if (t.kind == T_STRING && t.s.find(' ') != std::string::npos) {
        // "<type> <name>" header, then a bracketed array or one bare value
        ++i;
        Param p;
        const size_t sp = t.s.find_first_of(" \t\n\r\f\v");
        p.type = trim(t.s.substr(0, sp));
        p.name = sp == std::string::npos ? "" : trim(t.s.substr(sp + 1));
        if (p.type.empty() || p.name.empty())
          parse_err(t.pos, "parameter header " + dbg(t.s) + " must have form \"<type> <name>\"");
        p.pos = t.pos;
        const bool bracketed = i < toks.size() && toks[i].kind == T_LBRACKET;
        if (bracketed) ++i;
        int seen = -1;  // 0 floats, 1 ints, 2 strings
        for (;;) {
          if (i >= toks.size()) {
            if (bracketed) parse_err(last, "unexpected end of input");
            break;
          }
          const Token& v = toks[i];
          if (v.kind == T_RBRACKET && bracketed) {
            ++i;
            break;
          }
          if (v.kind == T_FLOAT) {
            if (seen == 2) parse_err(v.pos, "mixed types in parameter array");
            if (seen == 1) {
              for (int64_t x : p.ints) p.floats.push_back((double)x);
              p.ints.clear();
            }
            seen = 0;
            p.floats.push_back(v.f);
            ++i;
          } else if (v.kind == T_INT) {
            if (seen == 2) parse_err(v.pos, "mixed types in parameter array");
            if (seen == 0) {
              p.floats.push_back((double)v.i);
            } else {
              seen = 1;
              p.ints.push_back(v.i);
            }
            ++i;
          } else if (v.kind == T_STRING) {
            if (seen >= 0 && seen != 2) parse_err(v.pos, "mixed types in parameter array");
            seen = 2;
            p.strings.push_back(v.s);
            ++i;
          } else {
            if (bracketed) expected(v.pos, "value or `]`", tok_kind(v.kind));
            break;
          }
          if (!bracketed) break;
        }
        if (seen < 0) expected(t.pos, "parameter value", "(empty)");
        p.kind = seen == 0 ? V_FLOATS : seen == 1 ? V_INTS : V_STRINGS;
        if (p.kind == V_INTS && is_float_type(p.type)) {  // "rgb L" [15 15 15]
          for (int64_t x : p.ints) p.floats.push_back((double)x);
          p.ints.clear();
          p.kind = V_FLOATS;
        }
        d.params.push_back(std::move(p));
      } else if (t.kind == T_STRING) {
        Positional a;
        a.kind = P_STRING;
        a.s = t.s;
        d.positional.push_back(std::move(a));
        ++i;
      } else if (t.kind == T_INT || t.kind == T_FLOAT) {
        Positional a;
        a.kind = P_FLOAT;
        a.f = t.kind == T_INT ? (double)t.i : t.f;
        d.positional.push_back(std::move(a));
        ++i;
      } else if (t.kind == T_LBRACKET) {  // Transform / ConcatTransform arrays
        ++i;
        Positional a;
        a.kind = P_ARRAY;
        for (;;) {
          if (i >= toks.size()) parse_err(last, "unexpected end of input");
          const Token& v = toks[i];
          if (v.kind == T_RBRACKET) {
            ++i;
            break;
          }
          if (v.kind != T_INT && v.kind != T_FLOAT) expected(v.pos, "number or `]`", tok_kind(v.kind));
          a.arr.push_back(v.kind == T_INT ? (double)v.i : v.f);
          ++i;
        }
        d.positional.push_back(std::move(a));
      } else {
        expected(t.pos, "directive arg", "`]`");
      }
    }
    out.push_back(std::move(d));
  }
  return out;
}

// --------------------------------------------------------- Radiance HDR --
// io/hdr.rs:24-50 over image 0.25.9's hdr decoder.

struct Rgbe {
  uint8_t c[3], e;
};

struct HdrImage {
  uint32_t w = 0, h = 0;
  std::vector<float> rgb;
};

[[noreturn]] void hdr_err(const std::string& what) { fail(PYRE_EINVAL, "decode: " + what); }

HdrImage decode_radiance(const std::vector<uint8_t>& b) {
  size_t i = 0;
  auto line = [&](std::string* s) {
    s->clear();
    while (i < b.size() && b[i] != '\n') s->push_back((char)b[i++]);
    if (i >= b.size()) hdr_err("unexpected end of header");
    ++i;
  };
  static const char sig[] = "#?RADIANCE";
  if (b.size() < 10 || std::memcmp(b.data(), sig, 10) != 0) hdr_err("Radiance HDR signature is invalid");
  std::string s;
  i = 10;
  line(&s);  // rest of the signature line
  for (;;) {  // header: KEY=value lines and comments, up to an empty line
    line(&s);
    if (s.empty()) break;
    if (s[0] == '#') continue;
    const size_t eq = s.find('=');
    if (eq == std::string::npos) continue;
    if (trim(s.substr(0, eq)) == "FORMAT" && trim(s.substr(eq + 1)) != "32-bit_rle_rgbe")
      hdr_err("unsupported pixel format " + trim(s.substr(eq + 1)));
  }
  line(&s);  // "-Y <height> +X <width>"
  char ya[8], xa[8];
  long long hh = -1, ww = -1;
  char tail;
  if (std::sscanf(s.c_str(), "%7s %lld %7s %lld %c", ya, &hh, xa, &ww, &tail) != 4 || std::strcmp(ya, "-Y") ||
      std::strcmp(xa, "+X"))
    hdr_err("unsupported dimensions line " + dbg(s) + " (only -Y H +X W)");
  if (hh <= 0 || ww <= 0 || hh > 65535 || ww > 65535 || hh * ww > (1ll << 28)) hdr_err("image dimensions out of range");
  HdrImage img;
  img.w = (uint32_t)ww;
  img.h = (uint32_t)hh;
  std::vector<Rgbe> row(img.w);
  auto byte = [&]() -> uint8_t {
    if (i >= b.size()) hdr_err("unexpected end of pixel data");
    return b[i++];
  };
  auto pixel = [&]() {
    Rgbe p;
    p.c[0] = byte();
    p.c[1] = byte();
    p.c[2] = byte();
    p.e = byte();
    return p;
  };
  img.rgb.resize(3 * (size_t)img.w * img.h);
  for (uint32_t y = 0; y < img.h; ++y) {
    const Rgbe fb = pixel();
    if (fb.c[0] == 2 && fb.c[1] == 2 && fb.c[2] < 128) {
      // per-component run-length planes: R, G, B, E
      for (int comp = 0; comp < 4; ++comp) {
        size_t pos = 0;
        while (pos < img.w) {
          const uint8_t rl = byte();
          if (rl <= 128) {
            if (pos + rl > img.w) hdr_err("wrong scanline length");
            for (uint8_t k = 0; k < rl; ++k) {
              const uint8_t v = byte();
              if (comp < 3)
                row[pos + k].c[comp] = v;
              else
                row[pos + k].e = v;
            }
            pos += rl;
          } else {
            const uint32_t n = rl - 128u;
            if (pos + n > img.w) hdr_err("wrong scanline length");
            const uint8_t v = byte();
            for (uint32_t k = 0; k < n; ++k) {
              if (comp < 3)
                row[pos + k].c[comp] = v;
              else
                row[pos + k].e = v;
            }
            pos += n;
          }
        }
      }
    } else {
      // flat pixels with the old run-length markers (1, 1, 1, n): repeat the
      // previous pixel n times, consecutive markers scale by 256
      auto marker = [](const Rgbe& p) { return p.c[0] == 1 && p.c[1] == 1 && p.c[2] == 1; };
      if (marker(fb)) hdr_err("first pixel of a scanline is a run-length marker");
      row[0] = fb;
      size_t x = 1, mult = 1;
      Rgbe prev = fb;
      while (x < img.w) {
        const Rgbe p = pixel();
        if (marker(p)) {
          const size_t rl = (size_t)p.e * mult;
          mult *= 256;
          if (x + rl > img.w) hdr_err("wrong scanline length");
          for (size_t k = 0; k < rl; ++k) row[x + k] = prev;
          x += rl;
        } else {
          mult = 1;
          prev = p;
          row[x++] = p;
        }
      }
    }
    float* dst = img.rgb.data() + 3 * (size_t)y * img.w;
    for (uint32_t x = 0; x < img.w; ++x) {
      const Rgbe& p = row[x];
      // Rgbe8Pixel::to_hdr: channel * 2^(e - 136), exact in f32
      const float s = p.e == 0 ? 0.0f : std::ldexp(1.0f, (int)p.e - 136);
      for (int k = 0; k < 3; ++k) dst[3 * x + k] = p.e == 0 ? 0.0f : s * (float)p.c[k];
    }
  }
  return img;
}

HdrImage load_hdri(const std::string& path) {
  const std::string e = lower_ext(path);
  if (e != "hdr" && e != "pic") fail(PYRE_EINVAL, "unsupported environment-map extension: " + e);
  return decode_radiance(read_bytes(path));
}

void set_env(pyre_preset* p, HdrImage&& img, float intensity) {
  p->env_w = img.w;
  p->env_h = img.h;
  p->env_rgb = std::move(img.rgb);
  p->env_intensity = intensity;
}

// --------------------------------------------------------- PBRT builder --
// builder.rs:66-661

[[noreturn]] void build_err(const std::string& what) { fail(PYRE_EINVAL, "build: " + what); }

[[noreturn]] void malformed(const std::string& what, uint32_t line, const std::string& why) {
  build_err("malformed " + what + " on line " + std::to_string(line) + ": " + why);
}

const Param* find_param(const Directive& d, const char* name) {
  for (const Param& p : d.params)
    if (p.name == name) return &p;
  return nullptr;
}

std::vector<float> expect_floats(const Directive& d, size_t n, const char* what) {
  std::vector<float> f;
  for (const Positional& a : d.positional)
    if (a.kind == P_FLOAT) f.push_back((float)a.f);
  if (f.size() < n)
    malformed(what, d.pos.line,
              "expected " + std::to_string(n) + " float positionals, got " + std::to_string(f.size()));
  return f;
}

std::string expect_first_string(const Directive& d, const char* what) {
  for (const Positional& a : d.positional)
    if (a.kind == P_STRING) return a.s;
  malformed(what, d.pos.line, "expected a positional string (type name)");
}

bool first_float(const Param* p, double* v) {
  if (!p) return false;
  if (p->kind == V_FLOATS && !p->floats.empty()) return *v = p->floats[0], true;
  if (p->kind == V_INTS && !p->ints.empty()) return *v = (double)p->ints[0], true;
  return false;
}

// Rust's f64 as i64: truncation, saturating, NaN -> 0.
int64_t f64_to_i64(double x) {
  if (std::isnan(x)) return 0;
  if (x >= 9223372036854775807.0) return INT64_MAX;
  if (x <= -9223372036854775808.0) return INT64_MIN;
  return (int64_t)x;
}

bool first_int(const Param* p, int64_t* v) {
  if (!p) return false;
  if (p->kind == V_INTS && !p->ints.empty()) return *v = p->ints[0], true;
  if (p->kind == V_FLOATS && !p->floats.empty()) return *v = f64_to_i64(p->floats[0]), true;
  return false;
}

bool param_float(const Directive& d, const char* name, float* v) {
  double x;
  if (!first_float(find_param(d, name), &x)) return false;
  *v = (float)x;
  return true;
}

bool param_rgb(const Directive& d, const char* name, V3* v) {
  const Param* p = find_param(d, name);
  if (!p || p->kind != V_FLOATS) return false;
  if (p->floats.size() >= 3) return *v = v3((float)p->floats[0], (float)p->floats[1], (float)p->floats[2]), true;
  if (p->floats.size() == 1) return *v = splat((float)p->floats[0]), true;
  return false;
}

const std::string* first_string(const Param* p) {
  if (!p || p->kind != V_STRINGS || p->strings.empty()) return nullptr;
  return &p->strings[0];
}

pyre_material lambert_mat(V3 c) { return pyre_material{PYRE_MAT_LAMBERTIAN, {c.x, c.y, c.z}, 0.0f, 0.0f, 0.0f}; }
pyre_material disney_mat(V3 c, float m, float r, float s) {
  return pyre_material{PYRE_MAT_DISNEY, {c.x, c.y, c.z}, m, r, s};
}

// builder.rs:306-361
pyre_material build_material(const std::string& kind, const Directive& d) {
  if (kind == "matte") {
    V3 kd = splat(0.5f);
    param_rgb(d, "Kd", &kd);
    return lambert_mat(kd);
  }
  if (kind == "plastic") {
    V3 kd = splat(0.5f), ks = splat(0.5f);
    float rough = 0.2f;
    param_rgb(d, "Kd", &kd);
    param_rgb(d, "Ks", &ks);
    param_float(d, "roughness", &rough);
    return disney_mat(kd, 0.0f, rough, (ks.x + ks.y + ks.z) / 3.0f);
  }
  if (kind == "metal" || kind == "conductor") {
    V3 base = v3(1.0f, 0.766f, 0.336f);
    if (!param_rgb(d, "Kr", &base)) param_rgb(d, "reflectance", &base);
    float rough = 0.1f;
    param_float(d, "roughness", &rough);
    return disney_mat(base, 1.0f, rough, 0.5f);
  }
  return lambert_mat(splat(0.5f));  // unsupported: matte grey
}

struct Attr {
  M4 ctm;
  int64_t material = -1;
  bool area = false;
  V3 emission = splat(0.0f);
};

uint32_t default_material(pyre_preset* p) {
  p->materials.push_back(lambert_mat(splat(0.5f)));
  return (uint32_t)p->materials.size() - 1;
}

// builder.rs:532-564
PresetMesh uv_sphere_mesh(float radius, uint32_t us, uint32_t vs, const M4& ctm) {
  PresetMesh m;
  for (uint32_t v = 0; v <= vs; ++v) {
    const float phi = kPi * (float)v / (float)vs;
    const float sp = sinf(phi), cp = cosf(phi);
    for (uint32_t u = 0; u <= us; ++u) {
      const float theta = kTau * (float)u / (float)us;
      const V3 n = v3(sp * cosf(theta), cp, sp * sinf(theta));
      const V3 p = m4_point(ctm, radius * n);
      const V3 w = normalize(m4_vector(ctm, n));
      m.pos.insert(m.pos.end(), {p.x, p.y, p.z});
      m.nrm.insert(m.nrm.end(), {w.x, w.y, w.z});
    }
  }
  const uint32_t row = us + 1;
  for (uint32_t v = 0; v < vs; ++v)
    for (uint32_t u = 0; u < us; ++u) {
      const uint32_t i00 = v * row + u, i01 = v * row + u + 1, i10 = (v + 1) * row + u, i11 = (v + 1) * row + u + 1;
      m.idx.insert(m.idx.end(), {i00, i10, i11, i00, i11, i01});
    }
  return m;
}

// An emissive shape that is not a quad, as an emissive mesh light (with
// PYRE_PBRT_MESH_EMITTERS): like the quad conversion, the shape becomes a
// light and not a primitive.
void add_mesh_emitter(const PresetMesh& m, const Attr& cur, pyre_preset* p) {
  PresetMeshLight l;
  l.pos = m.pos;
  l.idx = m.idx;
  l.emission[0] = cur.emission.x;
  l.emission[1] = cur.emission.y;
  l.emission[2] = cur.emission.z;
  p->mesh_lights.push_back(std::move(l));
}

void add_shape(const std::string& kind, const Directive& d, const Attr& cur, pyre_preset* p, bool mesh_emitters) {
  if (kind == "sphere") {
    float radius = 1.0f;
    param_float(d, "radius", &radius);
    PresetMesh m = uv_sphere_mesh(radius, 48, 24, cur.ctm);
    if (cur.area && mesh_emitters) {
      add_mesh_emitter(m, cur, p);
      return;
    }
    // an area light on a sphere is ignored (builder.rs:509-518)
    m.material = cur.material >= 0 ? (uint32_t)cur.material : default_material(p);
    p->meshes.push_back(std::move(m));
    return;
  }
  if (kind != "trianglemesh") return;  // other shapes are skipped
  const Param* P = find_param(d, "P");
  if (!P || P->kind != V_FLOATS) build_err("Shape \"trianglemesh\" requires parameter \"P\"");
  const Param* I = find_param(d, "indices");
  if (!I || I->kind != V_INTS) build_err("Shape \"trianglemesh\" requires parameter \"indices\"");
  if (P->floats.size() % 3) malformed("trianglemesh", d.pos.line, "P length not a multiple of 3");
  if (I->ints.size() % 3) malformed("trianglemesh", d.pos.line, "indices length not a multiple of 3");
  PresetMesh m;
  const size_t nv = P->floats.size() / 3;
  for (size_t k = 0; k < nv; ++k) {
    const V3 w = m4_point(cur.ctm, v3((float)P->floats[3 * k], (float)P->floats[3 * k + 1], (float)P->floats[3 * k + 2]));
    m.pos.insert(m.pos.end(), {w.x, w.y, w.z});
  }
  for (int64_t x : I->ints) {
    const uint32_t v = (uint32_t)(uint64_t)x;  // `i as u32`
    // the reference indexes the positions with it (a panic when out of range)
    if (v >= nv) malformed("trianglemesh", d.pos.line, "index " + std::to_string(x) + " out of range");
    m.idx.push_back(v);
  }
  if (const Param* N = find_param(d, "N")) {
    if (N->kind == V_FLOATS) {
      const size_t nn = N->floats.size() / 3;
      if (nn != nv) malformed("trianglemesh", d.pos.line, "N holds " + std::to_string(nn) + " normals for " +
                                                                 std::to_string(nv) + " points");
      for (size_t k = 0; k < nn; ++k) {
        const V3 w = normalize(
            m4_vector(cur.ctm, v3((float)N->floats[3 * k], (float)N->floats[3 * k + 1], (float)N->floats[3 * k + 2])));
        m.nrm.insert(m.nrm.end(), {w.x, w.y, w.z});
      }
    }
  }
  if (cur.area && nv == 4 && m.idx.size() == 6) {
    // an emissive quad becomes a DiffuseAreaQuadLight (builder.rs:460-478)
    pyre_quad_light q;
    const V3 p0 = v3(m.pos[0], m.pos[1], m.pos[2]);
    const V3 eu = v3(m.pos[3], m.pos[4], m.pos[5]) - p0, ev = v3(m.pos[9], m.pos[10], m.pos[11]) - p0;
    const float f[12] = {p0.x, p0.y, p0.z, eu.x, eu.y, eu.z, ev.x, ev.y, ev.z,
                         cur.emission.x, cur.emission.y, cur.emission.z};
    std::memcpy(q.p0, f, 3 * sizeof(float));
    std::memcpy(q.edge_u, f + 3, 3 * sizeof(float));
    std::memcpy(q.edge_v, f + 6, 3 * sizeof(float));
    std::memcpy(q.emission, f + 9, 3 * sizeof(float));
    p->lights.push_back(q);
    return;
  }
  if (cur.area && mesh_emitters && !m.idx.empty()) {
    add_mesh_emitter(m, cur, p);
    return;
  }
  m.material = cur.material >= 0 ? (uint32_t)cur.material : default_material(p);
  p->meshes.push_back(std::move(m));
}

void add_light_source(const std::string& kind, const Directive& d, const std::string& base, pyre_preset* p) {
  if (kind != "infinite") return;  // other light sources are skipped
  float intensity = 1.0f;
  param_float(d, "scale", &intensity);
  if (const std::string* map = first_string(find_param(d, "mapname"))) {
    try {
      set_env(p, load_hdri(join(base, *map)), intensity);
    } catch (const LoadError& e) {
      build_err("HDRI load failed: " + e.msg);
    }
    return;
  }
  V3 l = splat(1.0f);
  param_rgb(d, "L", &l);
  HdrImage c;  // HdriEnvironmentLight::constant: one pixel
  c.w = c.h = 1;
  c.rgb = {l.x, l.y, l.z};
  set_env(p, std::move(c), intensity);
}

void build_pbrt(const std::vector<Directive>& dirs, const std::string& base, pyre_preset* p, bool mesh_emitters) {
  std::map<std::string, uint32_t> named;
  std::vector<Attr> attr_stack;
  std::vector<M4> xform_stack;
  Attr cur;
  cur.ctm = m4_identity();
  bool have_lookat = false, have_fov = false, in_world = false;
  V3 eye = splat(0.0f), look = splat(0.0f), up = splat(0.0f);
  float fov = 0.0f;
  for (const Directive& d : dirs) {
    const std::string& k = d.keyword;
    if (k == "LookAt") {
      const std::vector<float> f = expect_floats(d, 9, "LookAt");
      eye = v3(f[0], f[1], f[2]);
      look = v3(f[3], f[4], f[5]);
      up = v3(f[6], f[7], f[8]);
      have_lookat = true;
    } else if (k == "Camera") {
      expect_first_string(d, "Camera");  // only "perspective" is meant; the fov is read regardless
      double v;
      have_fov = first_float(find_param(d, "fov"), &v);
      fov = have_fov ? (float)v : 0.0f;
    } else if (k == "Film") {
      int64_t w = 800, h = 600;
      first_int(find_param(d, "xresolution"), &w);
      first_int(find_param(d, "yresolution"), &h);
      p->view.has_film = 1;
      p->view.width = (uint32_t)(uint64_t)w;
      p->view.height = (uint32_t)(uint64_t)h;
    } else if (k == "Identity") {
      cur.ctm = m4_identity();
    } else if (k == "Translate") {
      const std::vector<float> f = expect_floats(d, 3, "Translate");
      cur.ctm = m4_mul(cur.ctm, m4_translation(f[0], f[1], f[2]));
    } else if (k == "Scale") {
      const std::vector<float> f = expect_floats(d, 3, "Scale");
      cur.ctm = m4_mul(cur.ctm, m4_scale(f[0], f[1], f[2]));
    } else if (k == "Rotate") {
      const std::vector<float> f = expect_floats(d, 4, "Rotate");
      const float angle = f[0] * (kPi / 180.0f);  // f32::to_radians
      const V3 axis = normalize(v3(f[1], f[2], f[3]));
      const float h = angle * 0.5f;  // Quat::from_axis_angle
      const float s = sinf(h), c = cosf(h);
      const V3 q = axis * s;
      cur.ctm = m4_mul(cur.ctm, m4_from_quat(q.x, q.y, q.z, c));
    } else if (k == "Transform" || k == "ConcatTransform") {
      if (d.positional.empty() || d.positional[0].kind != P_ARRAY || d.positional[0].arr.size() != 16)
        malformed(k, d.pos.line, "expected a 16-float array");
      M4 m;  // the 16 values are the matrix's columns (Mat4::from_cols_array)
      for (int j = 0; j < 16; ++j) m.c[j / 4][j % 4] = (float)d.positional[0].arr[j];
      cur.ctm = k == "Transform" ? m : m4_mul(cur.ctm, m);
    } else if (k == "AttributeBegin") {
      attr_stack.push_back(cur);
    } else if (k == "AttributeEnd") {
      if (attr_stack.empty())
        build_err("AttributeEnd without matching AttributeBegin (line " + std::to_string(d.pos.line) + ")");
      cur = attr_stack.back();
      attr_stack.pop_back();
    } else if (k == "TransformBegin") {
      xform_stack.push_back(cur.ctm);
    } else if (k == "TransformEnd") {
      if (xform_stack.empty())
        build_err("TransformEnd without matching TransformBegin (line " + std::to_string(d.pos.line) + ")");
      cur.ctm = xform_stack.back();
      xform_stack.pop_back();
    } else if (k == "WorldBegin") {
      in_world = true;
      cur.ctm = m4_identity();
    } else if (k == "WorldEnd") {
      in_world = false;
    } else if (k == "Material") {
      const std::string name = expect_first_string(d, "Material");
      p->materials.push_back(build_material(name, d));
      cur.material = (int64_t)p->materials.size() - 1;
    } else if (k == "MakeNamedMaterial") {
      const std::string name = expect_first_string(d, "MakeNamedMaterial");
      const std::string* type = first_string(find_param(d, "type"));
      p->materials.push_back(build_material(type ? *type : std::string("matte"), d));
      named[name] = (uint32_t)p->materials.size() - 1;
    } else if (k == "NamedMaterial") {
      const std::string name = expect_first_string(d, "NamedMaterial");
      auto it = named.find(name);
      if (it == named.end()) build_err("named material " + dbg(name) + " not declared");
      cur.material = it->second;
    } else if (k == "AreaLightSource") {
      expect_first_string(d, "AreaLightSource");  // "diffuse" is meant
      V3 l = splat(1.0f);
      param_rgb(d, "L", &l);
      cur.area = true;
      cur.emission = l;
    } else if (k == "LightSource") {
      add_light_source(expect_first_string(d, "LightSource"), d, base, p);
    } else if (k == "Shape") {
      if (!in_world) continue;  // shapes outside the world block are ignored
      add_shape(expect_first_string(d, "Shape"), d, cur, p, mesh_emitters);
    }
    // Sampler, Integrator, Accelerator, PixelFilter and every other
    // directive are skipped
  }
  if (have_lookat && have_fov) {
    p->view.has_camera = 1;
    const float f[9] = {eye.x, eye.y, eye.z, look.x, look.y, look.z, up.x, up.y, up.z};
    std::memcpy(p->view.origin, f, 3 * sizeof(float));
    std::memcpy(p->view.target, f + 3, 3 * sizeof(float));
    std::memcpy(p->view.up, f + 6, 3 * sizeof(float));
    p->view.vfov_deg = fov;
  }
}

// ------------------------------------------------------------------ JSON --

struct Json {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double n = 0.0;
  std::string s;
  std::vector<Json> a;
  std::vector<std::pair<std::string, Json>> o;
  const Json* get(const char* key) const {
    if (kind != OBJ) return nullptr;
    for (const auto& kv : o)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

[[noreturn]] void gltf_err(const std::string& what) { fail(PYRE_EINVAL, "gltf: " + what); }

struct JsonParser {
  const char* p;
  const char* e;
  int depth = 0;
  void ws() {
    while (p < e && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) ++p;
  }
  [[noreturn]] void bad() { gltf_err("invalid JSON"); }
  static void put_utf8(std::string& s, uint32_t c) {
    if (c < 0x80) {
      s += (char)c;
    } else if (c < 0x800) {
      s += (char)(0xC0 | (c >> 6));
      s += (char)(0x80 | (c & 63));
    } else if (c < 0x10000) {
      s += (char)(0xE0 | (c >> 12));
      s += (char)(0x80 | ((c >> 6) & 63));
      s += (char)(0x80 | (c & 63));
    } else {
      s += (char)(0xF0 | (c >> 18));
      s += (char)(0x80 | ((c >> 12) & 63));
      s += (char)(0x80 | ((c >> 6) & 63));
      s += (char)(0x80 | (c & 63));
    }
  }
  uint32_t hex4() {
    if (e - p < 4) bad();
    uint32_t v = 0;
    for (int k = 0; k < 4; ++k) {
      const char c = *p++;
      v <<= 4;
      if (c >= '0' && c <= '9')
        v |= (uint32_t)(c - '0');
      else if (c >= 'a' && c <= 'f')
        v |= (uint32_t)(c - 'a' + 10);
      else if (c >= 'A' && c <= 'F')
        v |= (uint32_t)(c - 'A' + 10);
      else
        bad();
    }
    return v;
  }
  std::string str() {
    if (p >= e || *p != '"') bad();
    ++p;
    std::string s;
    for (;;) {
      if (p >= e) bad();
      const char c = *p++;
      if (c == '"') break;
      if ((unsigned char)c < 0x20) bad();
      if (c != '\\') {
        s += c;
        continue;
      }
      if (p >= e) bad();
      const char x = *p++;
      switch (x) {
        case '"': s += '"'; break;
        case '\\': s += '\\'; break;
        case '/': s += '/'; break;
        case 'b': s += '\b'; break;
        case 'f': s += '\f'; break;
        case 'n': s += '\n'; break;
        case 'r': s += '\r'; break;
        case 't': s += '\t'; break;
        case 'u': {
          uint32_t cp = hex4();
          if (cp >= 0xD800 && cp < 0xDC00) {
            if (e - p < 6 || p[0] != '\\' || p[1] != 'u') bad();
            p += 2;
            const uint32_t lo = hex4();
            if (lo < 0xDC00 || lo >= 0xE000) bad();
            cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
          } else if (cp >= 0xDC00 && cp < 0xE000) {
            bad();
          }
          put_utf8(s, cp);
          break;
        }
        default: bad();
      }
    }
    return s;
  }
  Json value() {
    if (++depth > 256) bad();
    ws();
    if (p >= e) bad();
    Json j;
    if (*p == '{') {
      ++p;
      j.kind = Json::OBJ;
      ws();
      if (p < e && *p == '}') {
        ++p;
      } else {
        for (;;) {
          ws();
          std::string k = str();
          ws();
          if (p >= e || *p != ':') bad();
          ++p;
          j.o.emplace_back(std::move(k), value());
          ws();
          if (p < e && *p == ',') {
            ++p;
            continue;
          }
          if (p < e && *p == '}') {
            ++p;
            break;
          }
          bad();
        }
      }
    } else if (*p == '[') {
      ++p;
      j.kind = Json::ARR;
      ws();
      if (p < e && *p == ']') {
        ++p;
      } else {
        for (;;) {
          j.a.push_back(value());
          ws();
          if (p < e && *p == ',') {
            ++p;
            continue;
          }
          if (p < e && *p == ']') {
            ++p;
            break;
          }
          bad();
        }
      }
    } else if (*p == '"') {
      j.kind = Json::STR;
      j.s = str();
    } else if (e - p >= 4 && !std::strncmp(p, "true", 4)) {
      p += 4;
      j.kind = Json::BOOL;
      j.b = true;
    } else if (e - p >= 5 && !std::strncmp(p, "false", 5)) {
      p += 5;
      j.kind = Json::BOOL;
    } else if (e - p >= 4 && !std::strncmp(p, "null", 4)) {
      p += 4;
    } else {
      const char* s0 = p;
      if (p < e && *p == '-') ++p;
      while (p < e && ((*p >= '0' && *p <= '9') || *p == '.' || *p == 'e' || *p == 'E' || *p == '+' || *p == '-'))
        ++p;
      double v;
      if (p == s0 || !parse_f64(std::string(s0, p), &v)) bad();
      j.kind = Json::NUM;
      j.n = v;
    }
    --depth;
    return j;
  }
};

Json parse_json(const char* b, size_t n) {
  JsonParser jp{b, b + n};
  Json j = jp.value();
  jp.ws();
  if (jp.p != jp.e) gltf_err("invalid JSON (trailing data)");
  return j;
}

// ------------------------------------------------------------------ glTF --
// io/gltf.rs:20-92 (document, buffers, default scene, node walk)

double num(const Json* j, double dflt) { return j && j->kind == Json::NUM ? j->n : dflt; }

// A JSON number that must be a non-negative integer no larger than `max`
// (absent: dflt). Rejected before any cast, so no huge, negative or
// fractional value reaches size arithmetic.
size_t json_size(const Json* j, size_t dflt, size_t max, const char* what) {
  if (!j) return dflt;
  if (j->kind != Json::NUM || !(j->n >= 0.0) || j->n != std::floor(j->n) || j->n > (double)max)
    gltf_err(std::string("invalid ") + what);
  return (size_t)j->n;
}

size_t index_of(const Json* j, size_t count, const char* what) {
  if (!j || j->kind != Json::NUM || j->n < 0 || j->n != std::floor(j->n) || j->n >= (double)count)
    gltf_err(std::string("invalid ") + what + " index");
  return (size_t)j->n;
}

const std::vector<Json>& arr(const Json& doc, const char* key) {
  static const std::vector<Json> none;
  const Json* j = doc.get(key);
  return j && j->kind == Json::ARR ? j->a : none;
}

std::vector<uint8_t> base64(const std::string& s) {
  auto val = [](char c) -> int {
    if (c >= 'A' && c <= 'Z') return c - 'A';
    if (c >= 'a' && c <= 'z') return c - 'a' + 26;
    if (c >= '0' && c <= '9') return c - '0' + 52;
    if (c == '+') return 62;
    if (c == '/') return 63;
    return -1;
  };
  std::vector<uint8_t> out;
  uint32_t acc = 0;
  int bits = 0;
  for (char c : s) {
    if (c == '=') break;
    const int v = val(c);
    if (v < 0) gltf_err("invalid base64 buffer");
    acc = (acc << 6) | (uint32_t)v;
    bits += 6;
    if (bits >= 8) {
      bits -= 8;
      out.push_back((uint8_t)(acc >> bits));
    }
  }
  return out;
}

std::string percent_decode(const std::string& s) {
  std::string o;
  for (size_t i = 0; i < s.size(); ++i) {
    if (s[i] == '%' && i + 2 < s.size() && std::isxdigit((unsigned char)s[i + 1]) &&
        std::isxdigit((unsigned char)s[i + 2])) {
      o += (char)std::stoi(s.substr(i + 1, 2), nullptr, 16);
      i += 2;
    } else {
      o += s[i];
    }
  }
  return o;
}

struct Accessor {
  const uint8_t* data = nullptr;  // null: all zeros (no bufferView)
  size_t count = 0, stride = 0, elem = 0, comps = 0;
  uint32_t ctype = 0;
  bool normalized = false;
};

size_t comp_size(uint32_t t) {
  switch (t) {
    case 5120: case 5121: return 1;
    case 5122: case 5123: return 2;
    case 5125: case 5126: return 4;
    default: gltf_err("unknown accessor componentType " + std::to_string(t));
  }
}

size_t type_comps(const std::string& t) {
  if (t == "SCALAR") return 1;
  if (t == "VEC2") return 2;
  if (t == "VEC3") return 3;
  if (t == "VEC4" || t == "MAT2") return 4;
  if (t == "MAT3") return 9;
  if (t == "MAT4") return 16;
  gltf_err("unknown accessor type " + dbg(t));
}

struct Gltf {
  Json doc;
  std::vector<std::vector<uint8_t>> buffers;

  Accessor accessor(size_t i) const {
    const std::vector<Json>& acc = arr(doc, "accessors");
    if (i >= acc.size()) gltf_err("invalid accessor index");
    const Json& a = acc[i];
    Accessor r;
    if (a.get("sparse")) gltf_err("sparse accessors are not supported");
    r.ctype = (uint32_t)json_size(a.get("componentType"), 0, 0xFFFFFFFFu, "accessor componentType");
    const Json* ty = a.get("type");
    r.comps = type_comps(ty && ty->kind == Json::STR ? ty->s : "");
    r.elem = comp_size(r.ctype) * r.comps;
    r.count = json_size(a.get("count"), 0, 1ull << 40, "accessor count");
    const Json* nz = a.get("normalized");
    r.normalized = nz && nz->kind == Json::BOOL && nz->b;
    const Json* bvj = a.get("bufferView");
    if (!bvj) {  // all zeros
      if (r.count > (1u << 26)) gltf_err("accessor without bufferView is too large");
      return r;
    }
    const std::vector<Json>& views = arr(doc, "bufferViews");
    const Json& bv = views[index_of(bvj, views.size(), "bufferView")];
    const std::vector<uint8_t>& buf = buffers[index_of(bv.get("buffer"), buffers.size(), "buffer")];
    // every size is at most the buffer's, and bounds are checked by
    // subtraction so no sum or product can wrap
    const size_t view_off = json_size(bv.get("byteOffset"), 0, buf.size(), "bufferView byteOffset");
    const size_t view_len = json_size(bv.get("byteLength"), 0, buf.size(), "bufferView byteLength");
    if (view_len > buf.size() - view_off) gltf_err("bufferView exceeds its buffer");
    const size_t off = json_size(a.get("byteOffset"), 0, view_len, "accessor byteOffset");
    r.stride = json_size(bv.get("byteStride"), 0, 252, "bufferView byteStride");
    if (r.stride == 0) r.stride = r.elem;
    if (r.stride < r.elem) gltf_err("bufferView byteStride below its accessor's element size");
    if (r.count) {
      if (r.elem > view_len - off) gltf_err("accessor exceeds its bufferView");
      if (r.count - 1 > (view_len - off - r.elem) / r.stride) gltf_err("accessor exceeds its bufferView");
    }
    r.data = buf.data() + view_off + off;
    return r;
  }
};

template <class T>
T load_le(const uint8_t* p) {
  T v;
  std::memcpy(&v, p, sizeof v);
  return v;
}

std::vector<float> read_vec3_f32(const Accessor& a, const char* what) {
  if (a.comps != 3 || a.ctype != 5126) gltf_err(std::string(what) + " must be VEC3 float");
  std::vector<float> out(3 * a.count, 0.0f);
  if (a.data)
    for (size_t i = 0; i < a.count; ++i)
      for (int k = 0; k < 3; ++k) out[3 * i + k] = load_le<float>(a.data + i * a.stride + 4 * k);
  return out;
}

std::vector<uint32_t> read_indices(const Accessor& a) {
  if (a.comps != 1 || (a.ctype != 5121 && a.ctype != 5123 && a.ctype != 5125))
    gltf_err("indices must be unsigned byte / short / int scalars");
  std::vector<uint32_t> out(a.count, 0);
  if (a.data)
    for (size_t i = 0; i < a.count; ++i) {
      const uint8_t* q = a.data + i * a.stride;
      out[i] = a.ctype == 5121 ? q[0] : a.ctype == 5123 ? load_le<uint16_t>(q) : load_le<uint32_t>(q);
    }
  return out;
}

// The node's local matrix: "matrix" (column-major) or T * R * S.
M4 node_matrix(const Json& n) {
  const Json* m = n.get("matrix");
  if (m && m->kind == Json::ARR) {
    if (m->a.size() != 16) gltf_err("node matrix must have 16 numbers");
    M4 r;
    for (int j = 0; j < 16; ++j) r.c[j / 4][j % 4] = (float)num(&m->a[j], 0.0);
    return r;
  }
  float t[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, s[3] = {1, 1, 1};
  if (const Json* j = n.get("translation"))
    for (size_t k = 0; k < 3 && k < j->a.size(); ++k) t[k] = (float)num(&j->a[k], 0.0);
  if (const Json* j = n.get("rotation"))
    for (size_t k = 0; k < 4 && k < j->a.size(); ++k) q[k] = (float)num(&j->a[k], 0.0);
  if (const Json* j = n.get("scale"))
    for (size_t k = 0; k < 3 && k < j->a.size(); ++k) s[k] = (float)num(&j->a[k], 0.0);
  return m4_mul(m4_mul(m4_translation(t[0], t[1], t[2]), m4_from_quat(q[0], q[1], q[2], q[3])),
                m4_scale(s[0], s[1], s[2]));
}

// Mat3::from_mat4(x).inverse().transpose(): the columns y*z, z*x, x*y over
// the determinant (glam's inverse is their transpose).
void normal_matrix(const M4& x, V3 col[3]) {
  const V3 a = v3(x.c[0][0], x.c[0][1], x.c[0][2]), b = v3(x.c[1][0], x.c[1][1], x.c[1][2]),
           c = v3(x.c[2][0], x.c[2][1], x.c[2][2]);
  const V3 t0 = cross(b, c), t1 = cross(c, a), t2 = cross(a, b);
  const float inv_det = 1.0f / dot(c, t2);
  col[0] = t0 * inv_det;
  col[1] = t1 * inv_det;
  col[2] = t2 * inv_det;
}

void walk_node(const Gltf& g, size_t ni, const M4& parent, pyre_preset* p, std::vector<uint8_t>& seen, int depth) {
  // glTF nodes form a forest: a node reached twice (two parents, a cycle)
  // is an invalid document
  if (seen[ni]) gltf_err("node " + std::to_string(ni) + " has more than one parent");
  if (depth > 4096) gltf_err("node hierarchy deeper than 4096");
  seen[ni] = 1;
  const std::vector<Json>& nodes = arr(g.doc, "nodes");
  const Json& n = nodes[ni];
  const M4 xf = m4_mul(parent, node_matrix(n));
  if (const Json* mj = n.get("mesh")) {
    const std::vector<Json>& meshes = arr(g.doc, "meshes");
    const Json& mesh = meshes[index_of(mj, meshes.size(), "mesh")];
    const std::vector<Json>& accs = arr(g.doc, "accessors");
    for (const Json& prim : arr(mesh, "primitives")) {
      const double mode = num(prim.get("mode"), 4.0);
      if (mode != 4.0) gltf_err("non-triangle primitive (mode: " + std::to_string((int)mode) + ")");
      const Json* attrs = prim.get("attributes");
      const Json* pa = attrs ? attrs->get("POSITION") : nullptr;
      if (!pa) gltf_err("primitive has no positions attribute");
      PresetMesh m;
      const std::vector<float> pos = read_vec3_f32(g.accessor(index_of(pa, accs.size(), "accessor")), "POSITION");
      const size_t nv = pos.size() / 3;
      for (size_t k = 0; k < nv; ++k) {
        const V3 w = m4_point(xf, v3(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]));
        m.pos.insert(m.pos.end(), {w.x, w.y, w.z});
      }
      if (const Json* ij = prim.get("indices")) {
        m.idx = read_indices(g.accessor(index_of(ij, accs.size(), "accessor")));
      } else {
        for (size_t k = 0; k < nv; ++k) m.idx.push_back((uint32_t)k);
      }
      m.idx.resize(m.idx.size() / 3 * 3);  // a trailing partial triangle is no triangle (n_indices / 3)
      for (uint32_t v : m.idx)
        if (v >= nv) gltf_err("index out of range");
      if (const Json* nj = attrs->get("NORMAL")) {
        const std::vector<float> nrm = read_vec3_f32(g.accessor(index_of(nj, accs.size(), "accessor")), "NORMAL");
        if (nrm.size() != pos.size()) gltf_err("NORMAL count differs from POSITION count");
        V3 c[3];
        normal_matrix(xf, c);
        for (size_t k = 0; k < nv; ++k) {
          // Mat3 * Vec3, then normalize_or_zero
          // Mat3 * Vec3, then normalize_or_zero
          const V3 r = (c[0] * nrm[3 * k] + c[1] * nrm[3 * k + 1]) + c[2] * nrm[3 * k + 2];
          const float rcp = 1.0f / length(r);
          const V3 w = (std::isfinite(rcp) && rcp > 0.0f) ? r * rcp : splat(0.0f);
          m.nrm.insert(m.nrm.end(), {w.x, w.y, w.z});
        }
      }
      m.material = 0;
      p->meshes.push_back(std::move(m));
    }
  }
  if (const Json* ch = n.get("children"))
    for (const Json& c : ch->a) walk_node(g, index_of(&c, nodes.size(), "node"), xf, p, seen, depth + 1);
}

void load_gltf(const std::string& path, pyre_preset* p) {
  const std::vector<uint8_t> file = read_bytes(path);
  Gltf g;
  std::vector<uint8_t> bin;
  bool glb = file.size() >= 12 && !std::memcmp(file.data(), "glTF", 4);
  if (glb) {
    const uint32_t version = load_le<uint32_t>(file.data() + 4), total = load_le<uint32_t>(file.data() + 8);
    if (version != 2) gltf_err("unsupported GLB version " + std::to_string(version));
    if (total > file.size()) gltf_err("GLB length exceeds the file");
    size_t off = 12;
    bool have_json = false;
    while (off + 8 <= total) {
      const uint32_t len = load_le<uint32_t>(file.data() + off), type = load_le<uint32_t>(file.data() + off + 4);
      if (off + 8 + (size_t)len > total) gltf_err("GLB chunk exceeds the file");
      const uint8_t* d = file.data() + off + 8;
      if (type == 0x4E4F534Au && !have_json) {  // "JSON"
        g.doc = parse_json((const char*)d, len);
        have_json = true;
      } else if (type == 0x004E4942u && bin.empty()) {  // "BIN\0"
        bin.assign(d, d + len);
      }
      off += 8 + (size_t)len;
    }
    if (!have_json) gltf_err("GLB has no JSON chunk");
  } else {
    g.doc = parse_json((const char*)file.data(), file.size());
  }
  if (g.doc.kind != Json::OBJ) gltf_err("document is not an object");
  const std::string base = dir_of(path);
  const std::vector<Json>& bufs = arr(g.doc, "buffers");
  for (size_t i = 0; i < bufs.size(); ++i) {
    const Json* uri = bufs[i].get("uri");
    std::vector<uint8_t> data;
    if (!uri) {
      if (!glb || i != 0) gltf_err("buffer without uri outside a GLB");
      data = bin;
    } else if (uri->kind != Json::STR) {
      gltf_err("invalid buffer uri");
    } else if (uri->s.compare(0, 5, "data:") == 0) {
      const size_t comma = uri->s.find(',');
      if (comma == std::string::npos || uri->s.substr(0, comma).find(";base64") == std::string::npos)
        gltf_err("unsupported data URI");
      data = base64(uri->s.substr(comma + 1));
    } else {
      data = read_bytes(join(base, percent_decode(uri->s)));
    }
    const size_t len = json_size(bufs[i].get("byteLength"), 0, 1ull << 40, "buffer byteLength");
    if (data.size() < len) gltf_err("buffer shorter than its byteLength");
    g.buffers.push_back(std::move(data));
  }
  p->materials.push_back(lambert_mat(splat(0.73f)));  // main.rs:131-133
  const std::vector<Json>& scenes = arr(g.doc, "scenes");
  if (scenes.empty()) return;
  size_t si = 0;
  if (const Json* s = g.doc.get("scene")) si = index_of(s, scenes.size(), "scene");
  const std::vector<Json>& nodes = arr(g.doc, "nodes");
  std::vector<uint8_t> seen(nodes.size(), 0);
  for (const Json& n : arr(scenes[si], "nodes"))
    walk_node(g, index_of(&n, nodes.size(), "node"), m4_identity(), p, seen, 0);
}

void put_err(const LoadError& e, char* err, size_t err_len) {
  if (err && err_len) {
    std::strncpy(err, e.msg.c_str(), err_len - 1);
    err[err_len - 1] = '\0';
  }
}

template <class F>
int guarded(char* err, size_t err_len, F&& f) {
  if (err && err_len) err[0] = '\0';
  try {
    f();
    return PYRE_OK;
  } catch (const LoadError& e) {
    put_err(e, err, err_len);
    return e.code;
  } catch (const std::bad_alloc&) {
    put_err(LoadError{PYRE_ENOMEM, "out of host memory"}, err, err_len);
    return PYRE_ENOMEM;
  }
}

}  // namespace

extern "C" {

int pyre_preset_load_pbrt(const char* path, pyre_preset** out, char* err, size_t err_len) {
  return pyre_preset_load_pbrt_ex(path, 0u, out, err, err_len);
}

int pyre_preset_load_pbrt_ex(const char* path, uint32_t flags, pyre_preset** out, char* err, size_t err_len) {
  if (!path || !out) return PYRE_EINVAL;
  if (flags & ~(uint32_t)PYRE_PBRT_MESH_EMITTERS) return PYRE_EINVAL;
  *out = nullptr;
  return guarded(err, err_len, [&] {
    std::unique_ptr<pyre_preset> p(new pyre_preset());
    p->kind = PRESET_LOADED;
    build_pbrt(parse(tokenize(read_bytes(path))), dir_of(path), p.get(), (flags & PYRE_PBRT_MESH_EMITTERS) != 0);
    preset_scene_bounds(p.get());
    preset_finish(p.get());
    *out = p.release();
  });
}

int pyre_preset_load_gltf(const char* path, pyre_preset** out, char* err, size_t err_len) {
  if (!path || !out) return PYRE_EINVAL;
  *out = nullptr;
  return guarded(err, err_len, [&] {
    std::unique_ptr<pyre_preset> p(new pyre_preset());
    p->kind = PRESET_LOADED;
    load_gltf(path, p.get());
    preset_scene_bounds(p.get());
    preset_finish(p.get());
    *out = p.release();
  });
}

int pyre_load_hdr(const char* path, uint32_t* width, uint32_t* height, float* rgb, size_t rgb_len, char* err,
                  size_t err_len) {
  if (!path || !width || !height) return PYRE_EINVAL;
  return guarded(err, err_len, [&] {
    HdrImage img = load_hdri(path);
    *width = img.w;
    *height = img.h;
    if (rgb) {
      if (rgb_len < img.rgb.size()) fail(PYRE_EINVAL, "rgb holds fewer than 3 * width * height floats");
      std::memcpy(rgb, img.rgb.data(), img.rgb.size() * sizeof(float));
    }
  });
}

int pyre_preset_set_env_hdr(pyre_preset* p, const char* path, float intensity, char* err, size_t err_len) {
  if (!p || !path) return PYRE_EINVAL;
  return guarded(err, err_len, [&] {
    set_env(p, load_hdri(path), intensity);
    preset_finish(p);
  });
}

}  // extern "C"
