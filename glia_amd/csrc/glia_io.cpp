// glia_io.cpp -- NetCDF classic input/output, segmentation split and config parsing for the RD
// path (include/glia_io.h).  Host code only; no external library.
//
// NetCDF classic layout (CDF-1 / CDF-2, all integers big-endian):
//   magic "CDF" version(1|2), numrecs, dim_list, gatt_list, var_list, then the data.
//   list = tag (0x0A dims, 0x0C attributes, 0x0B variables) + count, or 8 zero bytes if absent.
//   name = length + bytes padded to 4; attribute = name type count values (padded to 4);
//   variable = name ndims dimids[] vatt_list type vsize begin (4-byte begin in CDF-1, 8 in CDF-2).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/glia_io.h"
#include "../../include/glia_opt.h"
#include "../../include/glia_rd.h"

namespace {

struct Reader {
  const std::vector<unsigned char>& b;
  size_t p = 0;
  bool ok = true;
  explicit Reader(const std::vector<unsigned char>& buf) : b(buf) {}
  uint64_t be(int nbytes) {
    if (p + nbytes > b.size()) {
      ok = false;
      return 0;
    }
    uint64_t v = 0;
    for (int i = 0; i < nbytes; i++) v = (v << 8) | b[p + i];
    p += nbytes;
    return v;
  }
  uint32_t u32() { return (uint32_t)be(4); }
  void skip(size_t n) {
    if (p + n > b.size()) ok = false;
    p += n;
  }
  std::string name() {
    const uint32_t len = u32();
    if (!ok || p + len > b.size()) {
      ok = false;
      return std::string();
    }
    std::string s(reinterpret_cast<const char*>(b.data() + p), len);
    skip((len + 3u) & ~3u);
    return s;
  }
};

int type_size(int t) {
  switch (t) {
    case 1: case 2: return 1;
    case 3: return 2;
    case 4: case 5: return 4;
    case 6: return 8;
    default: return 0;
  }
}

struct NcVar {
  std::string name;
  std::vector<int64_t> dims;
  int type = 0;
  uint64_t begin = 0;
};

// parse the header; fills the variable list
bool parse_header(const std::vector<unsigned char>& buf, std::vector<NcVar>& vars) {
  Reader r(buf);
  if (buf.size() < 4 || buf[0] != 'C' || buf[1] != 'D' || buf[2] != 'F') return false;
  const int version = buf[3];
  if (version != 1 && version != 2) return false;  // CDF-5 (64-bit data) is not read
  r.skip(4);
  r.u32();  // numrecs
  std::vector<int64_t> dimlen;
  // dimensions
  uint32_t tag = r.u32(), n = r.u32();
  if (tag == 0x0A) {
    for (uint32_t i = 0; i < n && r.ok; i++) {
      r.name();
      dimlen.push_back(r.u32());
    }
  } else if (tag != 0 || n != 0) {
    return false;
  }
  auto skip_atts = [&]() {
    uint32_t t = r.u32(), na = r.u32();
    if (t == 0 && na == 0) return;
    if (t != 0x0C) {
      r.ok = false;
      return;
    }
    for (uint32_t i = 0; i < na && r.ok; i++) {
      r.name();
      const int at = (int)r.u32();
      const uint32_t cnt = r.u32();
      const int sz = type_size(at);
      if (!sz) {
        r.ok = false;
        return;
      }
      r.skip(((size_t)cnt * sz + 3u) & ~size_t(3));
    }
  };
  skip_atts();  // global attributes
  tag = r.u32();
  n = r.u32();
  if (tag == 0 && n == 0) return r.ok;
  if (tag != 0x0B) return false;
  for (uint32_t i = 0; i < n && r.ok; i++) {
    NcVar v;
    v.name = r.name();
    const uint32_t nd = r.u32();
    for (uint32_t k = 0; k < nd && r.ok; k++) {
      const uint32_t id = r.u32();
      if (id >= dimlen.size()) return false;
      v.dims.push_back(dimlen[id]);
    }
    skip_atts();
    v.type = (int)r.u32();
    r.u32();  // vsize
    v.begin = (version == 1) ? r.u32() : r.be(8);
    vars.push_back(v);
  }
  return r.ok;
}

bool load_file(const char* path, std::vector<unsigned char>& buf) {
  if (!path) return false;
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  buf.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return true;
}

const NcVar* find_var(const std::vector<NcVar>& vars, const char* name) {
  for (const auto& v : vars)
    if (v.name == (name ? name : "data")) return &v;
  return nullptr;
}

template <typename T>
int read_var(const char* path, const char* var, T* out, int64_t count) {
  std::vector<unsigned char> buf;
  std::vector<NcVar> vars;
  if (!out || !load_file(path, buf) || !parse_header(buf, vars)) return GLIA_RD_ERR_ARG;
  const NcVar* v = find_var(vars, var);
  if (!v) return GLIA_RD_ERR_ARG;
  int64_t total = 1;
  for (auto d : v->dims) total *= d;
  const int sz = type_size(v->type);
  if (total != count || !sz || v->type == 2) return GLIA_RD_ERR_ARG;
  if (v->begin + (uint64_t)total * sz > buf.size()) return GLIA_RD_ERR_ARG;
  const unsigned char* p = buf.data() + v->begin;
  for (int64_t i = 0; i < total; i++, p += sz) {
    uint64_t u = 0;
    for (int k = 0; k < sz; k++) u = (u << 8) | p[k];
    double x;
    switch (v->type) {
      case 1: x = (double)(int8_t)u; break;
      case 3: x = (double)(int16_t)u; break;
      case 4: x = (double)(int32_t)u; break;
      case 5: {
        const uint32_t w = (uint32_t)u;
        float f;
        std::memcpy(&f, &w, 4);
        x = f;
        break;
      }
      default: {
        std::memcpy(&x, &u, 8);
        break;
      }
    }
    out[i] = (T)x;
  }
  return 0;
}

void put_be(std::vector<unsigned char>& b, uint64_t v, int nbytes) {
  for (int i = nbytes - 1; i >= 0; i--) b.push_back((unsigned char)(v >> (8 * i)));
}
void put_name(std::vector<unsigned char>& b, const std::string& s) {
  put_be(b, s.size(), 4);
  b.insert(b.end(), s.begin(), s.end());
  while (b.size() % 4) b.push_back(0);
}

}  // namespace

extern "C" {

int glia_nc_inquire(const char* path, const char* var, int* ndims, int64_t* dims, int* type) {
  std::vector<unsigned char> buf;
  std::vector<NcVar> vars;
  if (!ndims || !dims || !type || !load_file(path, buf) || !parse_header(buf, vars)) return GLIA_RD_ERR_ARG;
  const NcVar* v = find_var(vars, var);
  if (!v || v->dims.size() > 8) return GLIA_RD_ERR_ARG;
  *ndims = (int)v->dims.size();
  for (size_t i = 0; i < v->dims.size(); i++) dims[i] = v->dims[i];
  *type = v->type;
  return 0;
}

int glia_nc_read_f32(const char* path, const char* var, float* out, int64_t count) {
  return read_var<float>(path, var, out, count);
}

int glia_nc_read_f64(const char* path, const char* var, double* out, int64_t count) {
  return read_var<double>(path, var, out, count);
}

int glia_nc_write3_f32(const char* path, const char* var, const int* n, const float* data, int as_double) {
  if (!path || !n || !data || n[0] <= 0 || n[1] <= 0 || n[2] <= 0) return GLIA_RD_ERR_ARG;
  const std::string vname = var ? var : "data";
  const int64_t total = (int64_t)n[0] * n[1] * n[2];
  const int sz = as_double ? 8 : 4;
  std::vector<unsigned char> h;
  h.insert(h.end(), {'C', 'D', 'F', 2});
  put_be(h, 0, 4);  // numrecs
  put_be(h, 0x0A, 4);
  put_be(h, 3, 4);
  const char* dn[3] = {"x", "y", "z"};
  for (int a = 0; a < 3; a++) {
    put_name(h, dn[a]);
    put_be(h, (uint64_t)n[a], 4);
  }
  put_be(h, 0x0C, 4);  // global attributes: "CDF-5 mode" = 0 (NC_INT)
  put_be(h, 1, 4);
  put_name(h, "CDF-5 mode");
  put_be(h, 4, 4);
  put_be(h, 1, 4);
  put_be(h, 0, 4);
  put_be(h, 0x0B, 4);  // one variable
  put_be(h, 1, 4);
  put_name(h, vname);
  put_be(h, 3, 4);
  for (int a = 0; a < 3; a++) put_be(h, (uint64_t)a, 4);
  put_be(h, 0, 4);  // no variable attributes
  put_be(h, 0, 4);
  put_be(h, as_double ? 6 : 5, 4);
  const uint64_t vsize = (uint64_t)total * sz;
  put_be(h, vsize > 0xFFFFFFFCull ? 0xFFFFFFFFull : ((vsize + 3) & ~3ull), 4);
  const uint64_t begin = h.size() + 8;
  put_be(h, begin, 8);
  std::vector<unsigned char> out;
  out.reserve(h.size() + vsize);
  out.insert(out.end(), h.begin(), h.end());
  for (int64_t i = 0; i < total; i++) {
    uint64_t u;
    if (as_double) {
      const double x = data[i];
      std::memcpy(&u, &x, 8);
    } else {
      uint32_t w;
      std::memcpy(&w, &data[i], 4);
      u = w;
    }
    put_be(out, u, sz);
  }
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) return GLIA_RD_ERR_ARG;
  f.write(reinterpret_cast<const char*>(out.data()), (std::streamsize)out.size());
  return f ? 0 : GLIA_RD_ERR_ARG;
}

int glia_edema_obs_mask(const float* tc, const float* ed, int64_t count, double lambda, float* mask) {
  if (!tc || !ed || !mask || count <= 0) return GLIA_RD_ERR_ARG;
  const float lam = (float)lambda;
  for (int64_t i = 0; i < count; i++) mask[i] = (tc[i] > 0.99f) ? 1.f : ((ed[i] > 0.99f) ? 0.f : lam);
  return 0;
}

int glia_phi_mesh(int np, const double cm[3], double sigma, double spacing_factor, double* centers) {
  if (np < 1 || !cm || !centers) return 0;
  const double sp = spacing_factor * sigma;
  if (np == 1) {
    centers[0] = cm[0], centers[1] = cm[1], centers[2] = cm[2];
    return 1;
  }
  const int h = (int)std::lround(std::pow((double)np, 1.0 / 3.0));
  int k = 0;
  auto put = [&](double a, double b, double c) {
    if (k >= np) return;
    centers[3 * k + 0] = a * sp + cm[0];
    centers[3 * k + 1] = b * sp + cm[1];
    centers[3 * k + 2] = c * sp + cm[2];
    k++;
  };
  if (np % 2) {
    const int e = (h - 1) / 2;
    for (int z = -e; z <= e; z++)
      for (int y = -e; y <= e; y++)
        for (int x = -e; x <= e; x++) put(x, y, z);
  } else {
    // the octant cubes: index i != 0 sits at i - sign(i) / 2 spacings
    const int e = h / 2;
    auto off = [](int i) { return i - (i > 0 ? 0.5 : -0.5); };
    for (int z = -e; z <= e; z++)
      for (int y = -e; y <= e; y++)
        for (int x = -e; x <= e; x++)
          if (x && y && z) put(off(x), off(y), off(z));
  }
  return k;
}

int glia_read_phi_mesh(const char* path, int max_np, double* centers, int* comps, int* np, double* sigma,
                       double* spacing) {
  if (!path || !np || max_np < 0 || (max_np > 0 && !centers)) return -2;
  FILE* f = std::fopen(path, "r");
  if (!f) return -1;
  std::string line;
  auto getline = [&](std::string& out) {
    out.clear();
    int ch;
    while ((ch = std::fgetc(f)) != EOF && ch != '\n') out.push_back((char)ch);
    return !(ch == EOF && out.empty());
  };
  double sg = 0, spc = 0;
  if (getline(line)) {
    // "sigma = s, spacing = d": the value after the first '=' up to the comma, then after the next '='
    const size_t e1 = line.find('='), c1 = line.find(',');
    if (e1 != std::string::npos && c1 != std::string::npos) sg = std::atof(line.c_str() + e1 + 1);
    const size_t e2 = c1 == std::string::npos ? std::string::npos : line.find('=', c1 + 1);
    if (e2 != std::string::npos) spc = std::atof(line.c_str() + e2 + 1);
  }
  getline(line);  // " centers = ["
  int k = 0;
  while (getline(line)) {
    if (line.find(']') != std::string::npos) break;
    size_t pos = 0;
    if (k < max_np) centers[3 * k] = centers[3 * k + 1] = centers[3 * k + 2] = 0;
    for (int t = 0;; t++) {
      const size_t comma = line.find(',', pos);
      const double v = std::atof(line.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos).c_str());
      if (k < max_np) {
        if (t < 3) centers[3 * k + t] = v;
        else if (t == 3 && comps) comps[k] = (int)v;
      }
      if (comma == std::string::npos) break;
      pos = comma + 1;
    }
    k++;
  }
  std::fclose(f);
  *np = k;
  if (sigma) *sigma = sg;
  if (spacing) *spacing = spc;
  return k > max_np ? -2 : 0;
}

int glia_read_concomp(const char* path, int max_comp, double* weights, double* centers, int* ncomp) {
  if (!path || !ncomp || max_comp < 0 || (max_comp > 0 && !weights)) return -2;
  std::ifstream f(path);
  if (!f) return -1;
  std::string line;
  std::getline(f, line);  // "#components"
  std::getline(f, line);
  const int nc = std::max(0, std::atoi(line.c_str()));
  std::getline(f, line);  // "center of mass:"
  for (int i = 0; i < nc; i++) {
    std::getline(f, line);
    size_t pos = 0;
    for (int t = 0; t < 3; t++) {
      const size_t comma = line.find(',', pos);
      const double v = std::atof(line.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos).c_str());
      if (centers && i < max_comp) centers[3 * i + t] = v;
      if (comma == std::string::npos) {
        for (int r = t + 1; r < 3; r++)
          if (centers && i < max_comp) centers[3 * i + r] = 0;
        break;
      }
      pos = comma + 1;
    }
  }
  std::getline(f, line);  // "relative mass:"
  for (int i = 0; i < nc; i++) {
    std::getline(f, line);
    if (i < max_comp) weights[i] = std::atof(line.c_str());
  }
  *ncomp = nc;
  return nc > max_comp ? -2 : 0;
}

int glia_split_segmentation(const float* seg, int64_t count, const int* labels, float* wm, float* gm, float* vt,
                            float* csf, float* tc, float* ed) {
  if (!seg || !labels || !wm || !gm || !vt || count <= 0) return GLIA_RD_ERR_ARG;
  const int wl = labels[0], gl = labels[1], vl = labels[2];
  if (wl <= 0 || gl <= 0 || vl <= 0) return GLIA_RD_ERR_ARG;  // WM, GM and VT are required
  const int cl = labels[3] > 0 ? labels[3] : -1, tl = labels[4] > 0 ? labels[4] : -1;
  const int nl = labels[5] > 0 ? labels[5] : -1, el = labels[6] > 0 ? labels[6] : -1;
  const int dl = labels[7] > 0 ? labels[7] : -1;
  for (int64_t i = 0; i < count; i++) {
    const float s = seg[i];
    wm[i] = (s == (float)wl) ? 1.f : 0.f;
    gm[i] = (s == (float)gl) ? 1.f : 0.f;
    vt[i] = (s == (float)vl) ? 1.f : 0.f;
    if (csf) csf[i] = (cl > 0 && s == (float)cl) ? 1.f : 0.f;
    if (tc) {
      float t = 0.f;
      if (tl > 0) {
        t = (s == (float)tl) ? 1.f : 0.f;
      } else if (nl > 0 && el > 0) {
        t = (s == (float)nl || s == (float)el) ? 1.f : 0.f;
      }
      tc[i] = t;
      if (t > 0) wm[i] = 1.f;  // the tumor core is treated as white matter
      if (dl > 0 && s == (float)dl) wm[i] = 1.f;  // and so is edema
    }
    // the edema map is filled only together with a tumor map, as in the reference
    if (ed) ed[i] = (tc && dl > 0 && s == (float)dl) ? 1.f : 0.f;
  }
  return 0;
}

int glia_rd_parse_config(const char* path, glia_rd_file_config* o) {
  if (!o) return GLIA_RD_ERR_ARG;
  std::memset(o, 0, sizeof(*o));
  // defaults of TumorParameters / OptimizerSettings (include/Parameters.h:15-235)
  o->n = 64;
  o->model = 1;
  o->nt = 1;
  o->dt = 0.5;
  o->k = 0.0;
  o->rho = 10.0;
  o->beta = 1e-4;
  o->smoothing_factor = 1.0;
  o->obs_threshold_1 = 0.0;
  o->np = 1;
  o->sigma_factor = 1.0;
  o->sigma_spacing = 1.5;
  for (int i = 0; i < 8; i++) o->atlas_labels[i] = o->patient_labels[i] = -1;
  o->obs_lambda = -1.0;
  o->smoothing_factor_data = 1.0;
  o->threshold_data_driven = 0.05;
  o->gaussian_volume_fraction = 0.0;
  o->sigma_data_driven = 2.0 * M_PI / 256;
  o->smoothing_factor_data_t0 = 1.0;
  o->obs_threshold_0 = 0.0;
  std::ifstream f(path ? path : "");
  if (!f) return GLIA_RD_ERR_ARG;
  std::string line;
  bool sigma_set = false;
  auto copy_path = [](char* dst, const std::string& v) {
    std::strncpy(dst, v.c_str(), 511);
    dst[511] = 0;
  };
  while (std::getline(f, line)) {
    line.erase(std::remove_if(line.begin(), line.end(), [](unsigned char c) { return std::isspace(c); }), line.end());
    if (line.empty() || line[0] == '#') continue;
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    const size_t eq = line.find('=');
    if (eq == std::string::npos) continue;
    const std::string k = line.substr(0, eq), v = line.substr(eq + 1);
    try {
      if (k == "n") o->n = std::stoi(v);
      else if (k == "model") o->model = std::stoi(v);
      else if (k == "nt_inv") o->nt = std::stoi(v);
      else if (k == "dt_inv") o->dt = std::stod(v);
      else if (k == "init_k") o->k = std::stod(v);
      else if (k == "init_rho") o->rho = std::stod(v);
      else if (k == "k_gm_wm") o->k_gm_wm = std::stod(v);
      else if (k == "r_gm_wm") o->r_gm_wm = std::stod(v);
      else if (k == "beta_p") o->beta = std::stod(v);
      else if (k == "smoothing_factor") o->smoothing_factor = std::stod(v);
      else if (k == "obs_threshold_1") o->obs_threshold_1 = std::stod(v);
      else if (k == "number_gaussians") o->np = std::stoi(v);
      else if (k == "sigma_factor") o->sigma_factor = std::stod(v), sigma_set = true;
      else if (k == "sigma_spacing") o->sigma_spacing = std::stod(v);
      else if (k == "store_phi") o->phi_store = std::stoi(v) > 0;
      else if (k == "store_adjoint") o->adjoint_store = std::stoi(v) > 0;
      else if (k == "a_seg_path") copy_path(o->a_seg_path, v);
      else if (k == "a_wm_path") copy_path(o->a_wm_path, v);
      else if (k == "a_gm_path") copy_path(o->a_gm_path, v);
      else if (k == "a_csf_path") copy_path(o->a_csf_path, v);
      else if (k == "a_vt_path") copy_path(o->a_vt_path, v);
      else if (k == "d1_path") copy_path(o->d1_path, v);
      else if (k == "obs_mask_path") copy_path(o->obs_mask_path, v);
      else if (k == "p_seg_path") copy_path(o->p_seg_path, v);
      else if (k == "obs_lambda") o->obs_lambda = std::stod(v);
      else if (k == "smoothing_factor_data") o->smoothing_factor_data = std::stod(v);
      else if (k == "obs_threshold_rel") o->obs_threshold_rel = std::stoi(v) > 0;
      else if (k == "threshold_data_driven") o->threshold_data_driven = std::stod(v);
      else if (k == "gaussian_volume_fraction") o->gaussian_volume_fraction = std::stod(v);
      else if (k == "support_data_path") copy_path(o->support_data_path, v);
      else if (k == "gaussian_cm_path") copy_path(o->gaussian_cm_path, v);
      else if (k == "data_comp_data_path") copy_path(o->data_comp_data_path, v);
      else if (k == "d0_path") copy_path(o->d0_path, v);
      else if (k == "smoothing_factor_data_t0") o->smoothing_factor_data_t0 = std::stod(v);
      else if (k == "obs_threshold_0") o->obs_threshold_0 = std::stod(v);
      else if (k == "atlas_labels" || k == "patient_labels") {
        int* dst = (k == "atlas_labels") ? o->atlas_labels : o->patient_labels;
        // [wm=6,gm=5,vt=7,csf=8,...] (src/utils/IO.cpp:126-170)
        std::string s = v;
        const size_t a = s.find('['), b = s.find(']');
        if (a != std::string::npos) s = s.substr(a + 1, (b == std::string::npos ? s.size() : b) - a - 1);
        const char* names[8] = {"wm", "gm", "vt", "csf", "tc", "nec", "en", "ed"};
        size_t pos = 0;
        while (pos <= s.size()) {
          size_t c = s.find(',', pos);
          if (c == std::string::npos) c = s.size();
          const std::string item = s.substr(pos, c - pos);
          const size_t e = item.find('=');
          if (e != std::string::npos)
            for (int i = 0; i < 8; i++)
              if (item.substr(0, e) == names[i]) dst[i] = std::stoi(item.substr(e + 1));
          pos = c + 1;
        }
      }
    } catch (...) {
      return GLIA_RD_ERR_ARG;
    }
  }
  // sigma_factor sets the data-driven sigma too (src/utils/IO.cpp:175-177), in units of 2 pi / n
  if (sigma_set && o->n > 0) o->sigma_data_driven = o->sigma_factor * 2.0 * M_PI / o->n;
  return 0;
}

// the optimizer keys of setParameter (src/utils/IO.cpp:82-100, 171-172) over the values the caller
// initialized; the same OptimizerSettings drive every solve of the reference, so they also go to
// the TIL and reaction/diffusion stages of the sparse-TIL settings
int glia_opt_parse_config(const char* path, glia_opt_settings* settings, glia_sparse_til_settings* sparse) {
  std::ifstream f(path ? path : "");
  if (!f) return GLIA_RD_ERR_ARG;
  std::vector<glia_opt_settings*> opts;
  if (settings) opts.push_back(settings);
  if (sparse) {
    opts.push_back(&sparse->til);
    opts.push_back(&sparse->rd);
  }
  std::string line;
  while (std::getline(f, line)) {
    line.erase(std::remove_if(line.begin(), line.end(), [](unsigned char c) { return std::isspace(c); }), line.end());
    const size_t hash = line.find('#');
    if (hash != std::string::npos) line = line.substr(0, hash);
    const size_t eq = line.find('=');
    if (line.empty() || eq == std::string::npos) continue;
    const std::string k = line.substr(0, eq), v = line.substr(eq + 1);
    try {
      for (glia_opt_settings* o : opts) {
        if (k == "newton_solver") o->newton_solver = (v == "GN") ? GLIA_OPT_GAUSS_NEWTON : GLIA_OPT_QUASI_NEWTON;
        else if (k == "line_search") o->line_search = (v == "armijo") ? GLIA_OPT_LS_ARMIJO : GLIA_OPT_LS_MORE_THUENTE;
        else if (k == "opttol_grad") o->gttol = std::stod(v);
        else if (k == "newton_maxit") o->max_it = std::stoi(v);
        else if (k == "krylov_maxit") o->krylov_maxit = std::stoi(v);
        else if (k == "lbfgs_vectors") o->lbfgs_vectors = std::stoi(v);
        else if (k == "ls_max_func_evals") o->ls_max_funcs = std::stoi(v);
      }
      if (sparse) {
        if (k == "gist_maxit") sparse->max_it = std::stoi(v);
        else if (k == "kappa_lb") sparse->k_lb = std::stod(v);
        else if (k == "kappa_ub") sparse->k_ub = std::stod(v);
        else if (k == "rho_lb") sparse->rho_lb = std::stod(v);
        else if (k == "rho_ub") sparse->rho_ub = std::stod(v);
        else if (k == "beta_p") sparse->beta = std::stod(v);
        else if (k == "invert_diff") sparse->diffusivity_inversion = std::stoi(v) > 0;
        else if (k == "invert_reac") sparse->reaction_inversion = std::stoi(v) > 0;
        else if (k == "sparsity_level") sparse->sparsity = std::stoi(v);
        else if (k == "thresh_component_weight") sparse->thresh_component_weight = std::stod(v);
      }
    } catch (...) {
      return GLIA_RD_ERR_ARG;
    }
  }
  return 0;
}

}  // extern "C"
