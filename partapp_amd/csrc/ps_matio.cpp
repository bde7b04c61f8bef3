// MAT-file (level 5) reader and writer for the files around the path:
// joint_<child>_<parent>.mat (objectdetect_learnparam.cpp:61-179), the
// cell_scoregrid / transform_Ti2 / transform_T2g score-grid files
// (partapp.cpp:792-857, matlab_cell_io.hpp:27-130) and the
// log_prob_grid / best_conf dumps (objectdetect_findrot.cpp:239-253,
// 1005-1011).  Written from the published level-5 layout: a 128-byte header
// (text, subsystem offset, version 0x0100, endian indicator "IM"), then data
// elements with an 8-byte tag (type, byte count) or a 4-byte "small element"
// tag, each padded to 8 bytes; miCOMPRESSED elements hold one zlib-deflated
// element; an miMATRIX element holds array flags (class), dimensions, name
// and the real part (in any numeric storage type), column-major.
// Little-endian files only (the endian indicator is checked).
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/partapp_ps.h"
#include "ps_plan.hpp"

namespace {

using ps::set_error;

enum : uint32_t {
  miINT8 = 1, miUINT8 = 2, miINT16 = 3, miUINT16 = 4, miINT32 = 5, miUINT32 = 6, miSINGLE = 7,
  miDOUBLE = 9, miINT64 = 12, miUINT64 = 13, miMATRIX = 14, miCOMPRESSED = 15, miUTF8 = 16
};
enum : int { mxCELL = 1, mxSTRUCT = 2, mxOBJECT = 3, mxCHAR = 4, mxSPARSE = 5, mxDOUBLE = 6, mxSINGLE = 7,
             mxINT8 = 8, mxUINT8 = 9, mxINT16 = 10, mxUINT16 = 11, mxINT32 = 12, mxUINT32 = 13, mxINT64 = 14,
             mxUINT64 = 15 };

struct Var {
  std::string name;
  int cls = 0;
  std::vector<long long> dims;  // MATLAB order
  std::vector<double> data;     // column-major
  std::vector<Var> cells;       // column-major
};

size_t type_size(uint32_t t) {
  switch (t) {
    case miINT8: case miUINT8: case miUTF8: return 1;
    case miINT16: case miUINT16: return 2;
    case miINT32: case miUINT32: case miSINGLE: return 4;
    case miDOUBLE: case miINT64: case miUINT64: return 8;
    default: return 0;
  }
}

template <class T>
T rd(const uint8_t *p) {
  T v;
  std::memcpy(&v, p, sizeof(T));
  return v;
}

bool to_doubles(uint32_t t, const uint8_t *p, size_t bytes, std::vector<double> &out) {
  const size_t sz = type_size(t);
  if (!sz || bytes % sz) return false;
  const size_t n = bytes / sz;
  out.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const uint8_t *q = p + i * sz;
    switch (t) {
      case miINT8: out[i] = rd<int8_t>(q); break;
      case miUINT8: case miUTF8: out[i] = rd<uint8_t>(q); break;
      case miINT16: out[i] = rd<int16_t>(q); break;
      case miUINT16: out[i] = rd<uint16_t>(q); break;
      case miINT32: out[i] = rd<int32_t>(q); break;
      case miUINT32: out[i] = rd<uint32_t>(q); break;
      case miSINGLE: out[i] = rd<float>(q); break;
      case miDOUBLE: out[i] = rd<double>(q); break;
      case miINT64: out[i] = (double)rd<int64_t>(q); break;
      case miUINT64: out[i] = (double)rd<uint64_t>(q); break;
    }
  }
  return true;
}

struct Reader {
  std::string err;
  // one data element at p (at most n bytes): type, payload, payload size, total
  // size including padding
  bool element(const uint8_t *p, size_t n, uint32_t &type, const uint8_t *&data, size_t &bytes, size_t &total) {
    if (n < 8) return fail("truncated element tag");
    const uint32_t t0 = rd<uint32_t>(p);
    if (t0 >> 16) {  // small element: 2-byte count, 2-byte type, data in the next 4 bytes
      type = t0 & 0xffff;
      bytes = t0 >> 16;
      if (bytes > 4) return fail("bad small element");
      data = p + 4;
      total = 8;
      return true;
    }
    type = t0;
    bytes = rd<uint32_t>(p + 4);
    data = p + 8;
    total = type == miCOMPRESSED ? 8 + bytes : 8 + ((bytes + 7) & ~size_t(7));
    if (8 + bytes > n) return fail("truncated element");
    if (total > n) total = n;  // the last element may be unpadded
    return true;
  }
  bool fail(const std::string &m) {
    if (err.empty()) err = m;
    return false;
  }

  bool matrix(const uint8_t *p, size_t n, Var &v) {
    size_t off = 0;
    uint32_t t;
    const uint8_t *d;
    size_t b, tot;
    // array flags
    if (!element(p + off, n - off, t, d, b, tot) || t != miUINT32 || b < 8) return fail("bad array flags");
    const uint32_t flags = rd<uint32_t>(d);
    v.cls = flags & 0xff;
    const bool complex = (flags & 0x800) != 0;
    off += tot;
    if (v.cls == mxSPARSE) return fail("sparse arrays are not supported");
    // dimensions
    if (!element(p + off, n - off, t, d, b, tot) || t != miINT32 || b < 8 || b % 4) return fail("bad dimensions");
    v.dims.clear();
    for (size_t i = 0; i < b / 4; ++i) v.dims.push_back(rd<int32_t>(d + 4 * i));
    off += tot;
    // name
    if (!element(p + off, n - off, t, d, b, tot) || (t != miINT8 && t != miUTF8)) return fail("bad array name");
    v.name.assign(reinterpret_cast<const char *>(d), b);
    off += tot;
    long long numel = 1;
    for (long long x : v.dims) numel *= x;
    if (v.cls == mxCELL) {
      v.cells.resize((size_t)numel);
      for (long long i = 0; i < numel; ++i) {
        if (!element(p + off, n - off, t, d, b, tot) || t != miMATRIX) return fail("bad cell element");
        if (b && !matrix(d, b, v.cells[(size_t)i])) return false;
        off += tot;
      }
      return true;
    }
    if (v.cls == mxSTRUCT || v.cls == mxOBJECT) return fail("struct / object arrays are not supported");
    if (numel == 0) return true;
    if (!element(p + off, n - off, t, d, b, tot)) return false;
    if (!to_doubles(t, d, b, v.data) || (long long)v.data.size() != numel) return fail("bad real part");
    (void)complex;  // the imaginary part, if any, is ignored
    return true;
  }

  bool file(const std::vector<uint8_t> &f, std::vector<Var> &vars) {
    if (f.size() < 128) return fail("not a MAT file (shorter than its header)");
    if (f[126] != 'I' || f[127] != 'M') return fail("not a little-endian level 5 MAT file");
    size_t off = 128;
    while (off + 8 <= f.size()) {
      uint32_t t;
      const uint8_t *d;
      size_t b, tot;
      if (!element(f.data() + off, f.size() - off, t, d, b, tot)) return false;
      if (t == miCOMPRESSED) {
        std::vector<uint8_t> raw;
        if (!inflate_all(d, b, raw)) return fail("corrupt compressed element");
        uint32_t t2;
        const uint8_t *d2;
        size_t b2, tot2;
        if (!element(raw.data(), raw.size(), t2, d2, b2, tot2)) return false;
        if (t2 == miMATRIX) {
          vars.emplace_back();
          if (!matrix(d2, b2, vars.back())) return false;
        }
      } else if (t == miMATRIX) {
        vars.emplace_back();
        if (!matrix(d, b, vars.back())) return false;
      }
      off += tot;
    }
    return true;
  }

  static bool inflate_all(const uint8_t *p, size_t n, std::vector<uint8_t> &out) {
    z_stream zs;
    std::memset(&zs, 0, sizeof(zs));
    if (inflateInit(&zs) != Z_OK) return false;
    zs.next_in = const_cast<Bytef *>(p);
    zs.avail_in = (uInt)n;
    out.resize(n * 4 + 1024);
    int st;
    do {
      if (zs.total_out == out.size()) out.resize(out.size() * 2);
      zs.next_out = out.data() + zs.total_out;
      zs.avail_out = (uInt)(out.size() - zs.total_out);
      st = inflate(&zs, Z_NO_FLUSH);
    } while (st == Z_OK);
    out.resize(zs.total_out);
    inflateEnd(&zs);
    return st == Z_STREAM_END;
  }
};

bool read_file(const char *path, std::vector<uint8_t> &buf) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  long sz = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  buf.resize(sz > 0 ? (size_t)sz : 0);
  bool ok = buf.empty() || std::fread(buf.data(), 1, buf.size(), f) == buf.size();
  std::fclose(f);
  return ok;
}

// ---------------------------------------------------------------- writing

void put_u32(std::vector<uint8_t> &o, uint32_t v) {
  uint8_t b[4];
  std::memcpy(b, &v, 4);
  o.insert(o.end(), b, b + 4);
}
void put_elem(std::vector<uint8_t> &o, uint32_t type, const void *data, size_t bytes) {
  put_u32(o, type);
  put_u32(o, (uint32_t)bytes);
  const uint8_t *p = static_cast<const uint8_t *>(data);
  o.insert(o.end(), p, p + bytes);
  while (o.size() % 8) o.push_back(0);
}

// one miMATRIX element of a real numeric array, data column-major
std::vector<uint8_t> matrix_element(const std::string &name, const std::vector<long long> &dims,
                                    const std::vector<double> &colmajor, bool single) {
  std::vector<uint8_t> body;
  const uint32_t flags[2] = {(uint32_t)(single ? mxSINGLE : mxDOUBLE), 0u};
  put_elem(body, miUINT32, flags, 8);
  std::vector<int32_t> d32(dims.begin(), dims.end());
  put_elem(body, miINT32, d32.data(), d32.size() * 4);
  put_elem(body, miINT8, name.data(), name.size());
  if (single) {
    std::vector<float> f(colmajor.begin(), colmajor.end());
    put_elem(body, miSINGLE, f.data(), f.size() * 4);
  } else {
    put_elem(body, miDOUBLE, colmajor.data(), colmajor.size() * 8);
  }
  std::vector<uint8_t> el;
  put_u32(el, miMATRIX);
  put_u32(el, (uint32_t)body.size());
  el.insert(el.end(), body.begin(), body.end());
  return el;
}

std::vector<uint8_t> header() {
  std::vector<uint8_t> h(128, ' ');
  const char *text = "MATLAB 5.0 MAT-file, written by partapp_ps";
  std::memcpy(h.data(), text, std::strlen(text));
  std::memset(h.data() + 116, 0, 8);  // no subsystem data
  h[124] = 0x00;
  h[125] = 0x01;  // version 0x0100, little-endian
  h[126] = 'I';
  h[127] = 'M';
  return h;
}

// column-major copy of a C-order array
std::vector<double> to_colmajor(const double *c, const std::vector<long long> &dims) {
  long long n = 1;
  for (long long d : dims) n *= d;
  std::vector<double> out((size_t)n);
  const int nd = (int)dims.size();
  std::vector<long long> idx(nd, 0);
  for (long long i = 0; i < n; ++i) {  // i walks C order
    long long cm = 0, stride = 1;
    for (int k = 0; k < nd; ++k) {
      cm += idx[k] * stride;
      stride *= dims[k];
    }
    out[(size_t)cm] = c[i];
    for (int k = nd - 1; k >= 0; --k) {
      if (++idx[k] < dims[k]) break;
      idx[k] = 0;
    }
  }
  return out;
}

}  // namespace

struct ps_mat {
  std::vector<Var> vars;
};

extern "C" int ps_mat_open(const char *path, ps_mat **out) {
  if (!path || !out) return set_error(PS_ERR_INVALID_ARG, "null argument");
  *out = nullptr;
  std::vector<uint8_t> buf;
  if (!read_file(path, buf)) return set_error(PS_ERR_IO, std::string("cannot read ") + path);
  ps_mat *m = new ps_mat();
  Reader r;
  if (!r.file(buf, m->vars)) {
    delete m;
    return set_error(PS_ERR_IO, std::string(path) + ": " + r.err);
  }
  *out = m;
  return PS_OK;
}

extern "C" void ps_mat_close(ps_mat *m) { delete m; }

extern "C" int ps_mat_var(const ps_mat *m, const char *name, long long cell_index, int *is_cell, int *ndim,
                          long long dims[8], double *out, long long capacity) {
  if (!m || !name || !ndim || !dims) return set_error(PS_ERR_INVALID_ARG, "null argument");
  const Var *v = nullptr;
  for (const Var &x : m->vars)
    if (x.name == name) {
      v = &x;
      break;
    }
  if (!v) return set_error(PS_ERR_NOT_FOUND, std::string("no variable ") + name);
  if (cell_index >= 0) {
    if (v->cls != mxCELL) return set_error(PS_ERR_INVALID_ARG, std::string(name) + " is not a cell array");
    if (cell_index >= (long long)v->cells.size()) return set_error(PS_ERR_INVALID_ARG, "cell index out of range");
    v = &v->cells[(size_t)cell_index];
  }
  if (is_cell) *is_cell = v->cls == mxCELL;
  if ((int)v->dims.size() > 8) return set_error(PS_ERR_UNSUPPORTED, "more than 8 dimensions");
  *ndim = (int)v->dims.size();
  long long n = 1;
  for (int k = 0; k < *ndim; ++k) {
    dims[k] = v->dims[k];
    n *= v->dims[k];
  }
  if (!out) return PS_OK;
  if (v->cls == mxCELL) return set_error(PS_ERR_INVALID_ARG, "a cell array has no numeric data; pass a cell index");
  if (capacity < n) return set_error(PS_ERR_INVALID_ARG, "output capacity too small");
  // column-major -> C order (the reference converts to c_storage_order on load)
  const int nd = *ndim;
  std::vector<long long> idx(nd, 0);
  for (long long i = 0; i < n; ++i) {
    long long cm = 0, stride = 1;
    for (int k = 0; k < nd; ++k) {
      cm += idx[k] * stride;
      stride *= v->dims[k];
    }
    out[i] = v->data[(size_t)cm];
    for (int k = nd - 1; k >= 0; --k) {
      if (++idx[k] < v->dims[k]) break;
      idx[k] = 0;
    }
  }
  return PS_OK;
}

extern "C" int ps_mat_write(const char *path, int append, const char *name, const double *data, int ndim,
                            const long long *dims, int single, int compress) {
  if (!path || !name || !dims || ndim < 1 || ndim > 8) return set_error(PS_ERR_INVALID_ARG, "bad argument");
  std::vector<long long> d(dims, dims + ndim);
  if (ndim == 1) d.push_back(1);  // MATLAB arrays have at least two dimensions: n x 1
  long long n = 1;
  for (long long x : d) {
    if (x < 0 || x > 0x7fffffff) return set_error(PS_ERR_INVALID_ARG, "bad dimension");
    n *= x;
  }
  if (n && !data) return set_error(PS_ERR_INVALID_ARG, "null data");
  std::vector<uint8_t> file;
  if (append) {
    if (!read_file(path, file) || file.size() < 128) file.clear();
  }
  if (file.empty()) file = header();
  std::vector<uint8_t> el = matrix_element(name, d, n ? to_colmajor(data, d) : std::vector<double>(), single != 0);
  if (compress) {
    uLongf zn = compressBound((uLong)el.size());
    std::vector<uint8_t> z(zn);
    if (compress2(z.data(), &zn, el.data(), (uLong)el.size(), Z_DEFAULT_COMPRESSION) != Z_OK)
      return set_error(PS_ERR_IO, "deflate failed");
    put_u32(file, miCOMPRESSED);
    put_u32(file, (uint32_t)zn);
    file.insert(file.end(), z.begin(), z.begin() + zn);
  } else {
    file.insert(file.end(), el.begin(), el.end());
  }
  FILE *f = std::fopen(path, "wb");
  if (!f) return set_error(PS_ERR_IO, std::string("cannot write ") + path);
  const bool ok = std::fwrite(file.data(), 1, file.size(), f) == file.size();
  std::fclose(f);
  if (!ok) return set_error(PS_ERR_IO, std::string("short write to ") + path);
  return PS_OK;
}

// load_joint (objectdetect_learnparam.cpp:126-170) then the per-joint part of
// loadJoints (objectdetect_aux.cpp:96-136): flip mirrors x (C -> T C T,
// offsets -> T v with T = diag(-1, 1), rot_mean negated for rotation joints),
// ids become 0-based, and C must have det > 0.
extern "C" int ps_load_joint_mat(const char *path, int flip, ps_joint *out, int *type) {
  if (!path || !out) return set_error(PS_ERR_INVALID_ARG, "null argument");
  ps_mat *m = nullptr;
  int st = ps_mat_open(path, &m);
  if (st != PS_OK) return st;
  auto scalar = [&](const char *nm, double &v) -> int {
    int nd;
    long long dims[8];
    int s = ps_mat_var(m, nm, -1, nullptr, &nd, dims, nullptr, 0);
    if (s != PS_OK) return s;
    if (nd != 2 || dims[0] != 1 || dims[1] != 1) return set_error(PS_ERR_INVALID_ARG, std::string(nm) + " is not 1x1");
    return ps_mat_var(m, nm, -1, nullptr, &nd, dims, &v, 1);
  };
  auto vec2 = [&](const char *nm, double v[2]) -> int {
    int nd;
    long long dims[8];
    int s = ps_mat_var(m, nm, -1, nullptr, &nd, dims, nullptr, 0);
    if (s != PS_OK) return s;
    if (nd != 2 || !((dims[0] == 2 && dims[1] == 1) || (dims[0] == 1 && dims[1] == 2)))
      return set_error(PS_ERR_INVALID_ARG, std::string(nm) + " is not a 2-vector");
    return ps_mat_var(m, nm, -1, nullptr, &nd, dims, v, 2);
  };
  double d_type = 0, d_child = 0, d_parent = 0, C[4];
  ps_joint j;
  std::memset(&j, 0, sizeof(j));
  st = scalar("type", d_type);
  if (st == PS_OK) st = scalar("child_idx", d_child);
  if (st == PS_OK) st = scalar("parent_idx", d_parent);
  if (st == PS_OK) st = vec2("offset_c", j.offset_c);
  if (st == PS_OK) st = vec2("offset_p", j.offset_p);
  if (st == PS_OK) {
    int nd;
    long long dims[8];
    st = ps_mat_var(m, "C", -1, nullptr, &nd, dims, nullptr, 0);
    if (st == PS_OK && (nd != 2 || dims[0] != 2 || dims[1] != 2)) st = set_error(PS_ERR_INVALID_ARG, "C is not 2x2");
    if (st == PS_OK) st = ps_mat_var(m, "C", -1, nullptr, &nd, dims, C, 4);
  }
  const int jtype = (int)d_type;
  j.rot_mean = 0.0;
  j.rot_sigma = 0.0;
  if (st == PS_OK && jtype == PS_JOINT_ROT_GAUSSIAN) {
    st = scalar("rot_sigma", j.rot_sigma);
    if (st == PS_OK) st = scalar("rot_mean", j.rot_mean);
  }
  ps_mat_close(m);
  if (st != PS_OK) return st;
  if (jtype != PS_JOINT_POS_GAUSSIAN && jtype != PS_JOINT_ROT_GAUSSIAN)
    return set_error(PS_ERR_INVALID_ARG, "unknown joint type " + std::to_string(jtype));
  if (flip) {
    // T C T with T = diag(-1, 1): the off-diagonal terms change sign
    C[1] = -C[1];
    C[2] = -C[2];
    j.offset_c[0] = -j.offset_c[0];
    j.offset_p[0] = -j.offset_p[0];
    if (jtype == PS_JOINT_ROT_GAUSSIAN) j.rot_mean = -j.rot_mean;
  }
  for (int i = 0; i < 4; ++i) j.C[i] = C[i];
  j.child_idx = (int)d_child - 1;
  j.parent_idx = (int)d_parent - 1;
  if (!(C[0] * C[3] - C[1] * C[2] > 0)) return set_error(PS_ERR_COVARIANCE, "det C must be > 0");
  *out = j;
  if (type) *type = jtype;
  return PS_OK;
}
