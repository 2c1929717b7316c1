// The reference's native entry point cufd (Src/libCUFD.cu:37-799) on top of
// the context API: reads the para and survey JSON files, the observed-data
// .bin files, and writes synthetic data and scratch files with the same names
// and byte layout as the reference. Errors are returned, never exit().
#include <hip/hip_runtime.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/efwi.h"
#include "efwi_host.h"

namespace efwi {
namespace {

// ---------------------------------------------------------------------------
// Minimal JSON reader (objects, arrays, numbers, strings, literals): enough
// for the files that fwi_utils.paraGen_3d / surveyGen_3d write with json.dump.
// ---------------------------------------------------------------------------
struct JVal {
  enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
  bool b = false;
  double num = 0.0;
  bool is_int = false;
  std::string str;
  std::vector<JVal> arr;
  std::map<std::string, JVal> obj;
  const JVal* get(const std::string& k) const {
    if (kind != OBJ) return nullptr;
    auto it = obj.find(k);
    return it == obj.end() ? nullptr : &it->second;
  }
};

struct JParser {
  static constexpr int kMaxDepth = 64;  // nesting of the files written by json.dump: 2
  const char* p;
  const char* end;
  std::string err;
  int depth = 0;
  void ws() {
    while (p < end && std::isspace((unsigned char)*p)) p++;
  }
  bool parse(JVal& v) {
    ws();
    if (p >= end) return fail("unexpected end of input");
    char c = *p;
    if (c == '{' || c == '[') {
      if (depth >= kMaxDepth) return fail("nesting deeper than " + std::to_string(kMaxDepth));
      depth++;
      const bool ok = c == '{' ? obj(v) : arr(v);
      depth--;
      return ok;
    }
    if (c == '"') {
      v.kind = JVal::STR;
      return string(v.str);
    }
    if (c == 't' && end - p >= 4 && !std::strncmp(p, "true", 4)) {
      v.kind = JVal::BOOL; v.b = true; p += 4; return true;
    }
    if (c == 'f' && end - p >= 5 && !std::strncmp(p, "false", 5)) {
      v.kind = JVal::BOOL; v.b = false; p += 5; return true;
    }
    if (c == 'n' && end - p >= 4 && !std::strncmp(p, "null", 4)) {
      v.kind = JVal::NUL; p += 4; return true;
    }
    return number(v);
  }
  bool fail(const std::string& m) {
    if (err.empty()) err = m;
    return false;
  }
  bool string(std::string& out) {
    p++;  // opening quote
    out.clear();
    while (p < end && *p != '"') {
      if (*p == '\\') {
        p++;
        if (p >= end) return fail("bad escape");
        char e = *p;
        if (e == 'n') out += '\n';
        else if (e == 't') out += '\t';
        else if (e == 'u') {  // keep \uXXXX as '?' -- file names in these files are ASCII
          if (end - p < 5) return fail("bad \\u escape");
          out += '?';
          p += 4;
        } else out += e;
        p++;
      } else {
        out += *p++;
      }
    }
    if (p >= end) return fail("unterminated string");
    p++;
    return true;
  }
  bool number(JVal& v) {
    char* q = nullptr;
    const char* s = p;
    double d = std::strtod(p, &q);
    if (q == p) return fail("bad value");
    bool frac = false;
    for (const char* t = s; t < q; t++)
      if (*t == '.' || *t == 'e' || *t == 'E') frac = true;
    v.kind = JVal::NUM;
    v.num = d;
    v.is_int = !frac;
    p = q;
    return true;
  }
  bool arr(JVal& v) {
    v.kind = JVal::ARR;
    p++;
    ws();
    if (p < end && *p == ']') { p++; return true; }
    while (true) {
      v.arr.emplace_back();
      if (!parse(v.arr.back())) return false;
      ws();
      if (p < end && *p == ',') { p++; continue; }
      if (p < end && *p == ']') { p++; return true; }
      return fail("expected , or ]");
    }
  }
  bool obj(JVal& v) {
    v.kind = JVal::OBJ;
    p++;
    ws();
    if (p < end && *p == '}') { p++; return true; }
    while (true) {
      ws();
      if (p >= end || *p != '"') return fail("expected key");
      std::string k;
      if (!string(k)) return false;
      ws();
      if (p >= end || *p != ':') return fail("expected :");
      p++;
      if (!parse(v.obj[k])) return false;
      ws();
      if (p < end && *p == ',') { p++; continue; }
      if (p < end && *p == '}') { p++; return true; }
      return fail("expected , or }");
    }
  }
};

int read_json(const std::string& fname, JVal& out) {
  FILE* f = std::fopen(fname.c_str(), "rb");
  if (!f) return fail(EFWI_ERR_IO, "cannot open " + fname);
  std::string text;
  char buf[1 << 16];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
  std::fclose(f);
  JParser P{text.data(), text.data() + text.size(), ""};
  if (!P.parse(out) || out.kind != JVal::OBJ)
    return fail(EFWI_ERR_IO, "malformed JSON in " + fname + (P.err.empty() ? "" : ": " + P.err));
  return EFWI_OK;
}

// an integer-valued number in int range (no silent truncation or overflow)
bool as_int(const JVal& x, int* v) {
  if (x.kind != JVal::NUM || !(x.num >= -2147483648.0 && x.num <= 2147483647.0) || x.num != std::floor(x.num))
    return false;
  *v = (int)x.num;
  return true;
}
bool get_int(const JVal& o, const char* k, int* v) {
  const JVal* x = o.get(k);
  return x && as_int(*x, v);
}
bool get_num(const JVal& o, const char* k, double* v) {
  const JVal* x = o.get(k);
  if (!x || x->kind != JVal::NUM) return false;
  *v = x->num;
  return true;
}
bool get_str(const JVal& o, const char* k, std::string* v) {
  const JVal* x = o.get(k);
  if (!x || x->kind != JVal::STR) return false;
  *v = x->str;
  return true;
}
bool get_ints(const JVal& o, const char* k, std::vector<int>& v) {
  const JVal* x = o.get(k);
  if (!x || x->kind != JVal::ARR) return false;
  v.clear();
  for (auto& e : x->arr) {
    int i;
    if (!as_int(e, &i)) return false;
    v.push_back(i);
  }
  return true;
}
bool get_floats(const JVal& o, const char* k, std::vector<float>& v) {
  const JVal* x = o.get(k);
  if (!x || x->kind != JVal::ARR) return false;
  v.clear();
  for (auto& e : x->arr) {
    if (e.kind != JVal::NUM) return false;
    v.push_back((float)e.num);
  }
  return true;
}

// The para file (Src/Parameter.cpp:41-184).
struct Para {
  int nz, nx, ny, nt, npml;
  float dz, dx, dy, dt, f0;
  std::string survey, data_dir, scratch_dir;
  bool save_scratch = false, if_win = false;
};

int read_para(const std::string& fname, Para& P) {
  JVal j;
  int rc = read_json(fname, j);
  if (rc) return rc;
  double d;
  auto need_i = [&](const char* k, int* v) { return get_int(j, k, v); };
  if (!need_i("nz", &P.nz) || !need_i("nx", &P.nx) || !need_i("ny", &P.ny) || !need_i("nSteps", &P.nt) ||
      !need_i("nPoints_pml", &P.npml))
    return fail(EFWI_ERR_IO, fname + ": missing one of nz, nx, ny, nSteps, nPoints_pml");
  if (!get_num(j, "dz", &d)) return fail(EFWI_ERR_IO, fname + ": missing dz");
  P.dz = (float)d;
  if (!get_num(j, "dx", &d)) return fail(EFWI_ERR_IO, fname + ": missing dx");
  P.dx = (float)d;
  if (!get_num(j, "dy", &d)) return fail(EFWI_ERR_IO, fname + ": missing dy");
  P.dy = (float)d;
  if (!get_num(j, "dt", &d)) return fail(EFWI_ERR_IO, fname + ": missing dt");
  P.dt = (float)d;
  if (!get_num(j, "f0", &d)) return fail(EFWI_ERR_IO, fname + ": missing f0");
  P.f0 = (float)d;
  if (!get_str(j, "survey_fname", &P.survey)) return fail(EFWI_ERR_IO, fname + ": missing survey_fname");
  if (!get_str(j, "data_dir_name", &P.data_dir)) return fail(EFWI_ERR_IO, fname + ": missing data_dir_name");
  if (get_str(j, "scratch_dir_name", &P.scratch_dir)) P.save_scratch = true;
  const JVal* w = j.get("if_win");
  if (w && w->kind == JVal::BOOL) P.if_win = w->b;
  return EFWI_OK;
}

int load_bin(const std::string& fname, float* dst, size_t n) {
  FILE* f = std::fopen(fname.c_str(), "rb");
  if (!f) return fail(EFWI_ERR_IO, "cannot open " + fname);
  size_t got = std::fread(dst, sizeof(float), n, f);
  std::fclose(f);
  if (got != n)
    return fail(EFWI_ERR_IO, fname + ": expected " + std::to_string(n) + " floats, read " + std::to_string(got));
  return EFWI_OK;
}

int write_bin(const std::string& fname, const float* src, size_t n) {
  FILE* f = std::fopen(fname.c_str(), "wb");
  if (!f) return fail(EFWI_ERR_IO, "cannot write " + fname);
  size_t put = std::fwrite(src, sizeof(float), n, f);
  std::fclose(f);
  if (put != n) return fail(EFWI_ERR_IO, "short write to " + fname);
  return EFWI_OK;
}

struct ShotDesc {
  int id;
  int sz, sx, sy;
  std::vector<int> rec;  // [nrec][3]
  std::vector<float> ws, we, wt;
  float srcw = 1.0f;
};

// One "shot<id>" object of the survey file (Src/Src_Rec.cu:74-206).
int read_shot(const JVal& sv, int id, int npml, bool if_win, ShotDesc& s) {
  const std::string key = "shot" + std::to_string(id);
  const JVal* o = sv.get(key);
  if (!o || o->kind != JVal::OBJ) return fail(EFWI_ERR_IO, "survey has no object " + key);
  int nrec;
  if (!get_int(*o, "z_src", &s.sz) || !get_int(*o, "x_src", &s.sx) || !get_int(*o, "y_src", &s.sy) ||
      !get_int(*o, "nrec", &nrec))
    return fail(EFWI_ERR_IO, key + ": missing z_src/x_src/y_src/nrec");
  if (nrec < 0) return fail(EFWI_ERR_IO, key + ": negative nrec");
  std::vector<int> zr, xr, yr;
  if (!get_ints(*o, "z_rec", zr) || !get_ints(*o, "x_rec", xr) || !get_ints(*o, "y_rec", yr))
    return fail(EFWI_ERR_IO, key + ": missing z_rec/x_rec/y_rec");
  if ((int)zr.size() < nrec || (int)xr.size() < nrec || (int)yr.size() < nrec)
    return fail(EFWI_ERR_IO, key + ": fewer receiver coordinates than nrec");
  s.id = id;
  s.sz += npml; s.sx += npml; s.sy += npml;
  s.rec.resize(3 * (size_t)nrec);
  for (int r = 0; r < nrec; r++) {
    s.rec[3 * r] = zr[r] + npml;
    s.rec[3 * r + 1] = xr[r] + npml;
    s.rec[3 * r + 2] = yr[r] + npml;
  }
  if (if_win) {
    if (!get_floats(*o, "win_start", s.ws) || !get_floats(*o, "win_end", s.we))
      return fail(EFWI_ERR_IO, key + ": if_win needs win_start and win_end");
    if ((int)s.ws.size() < nrec || (int)s.we.size() < nrec)
      return fail(EFWI_ERR_IO, key + ": window arrays shorter than nrec");
  }
  if (o->get("weights")) {
    if (!get_floats(*o, "weights", s.wt) || (int)s.wt.size() < nrec)
      return fail(EFWI_ERR_IO, key + ": bad weights");
  }
  double d;
  if (get_num(*o, "src_weight", &d)) s.srcw = (float)d;
  return EFWI_OK;
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) hipFree(p);
  }
};

}  // namespace
}  // namespace efwi

extern "C" int efwi_cufd(float* misfit, float* grad_lambda, float* grad_mu, float* grad_rho,
                         const float* lambda, const float* mu, const float* rho, const float* stf,
                         int calc_id, int gpu_id, int group_size, const int* shot_ids,
                         const char* para_fname) {
  using namespace efwi;
  if (calc_id < 0 || calc_id > 2) return fail(EFWI_ERR_ARG, "Invalid calc_id " + std::to_string(calc_id));
  if (!lambda || !mu || !rho || !stf || !para_fname || (group_size > 0 && !shot_ids))
    return fail(EFWI_ERR_ARG, "NULL argument");
  if (calc_id == EFWI_CALC_GRADIENT && (!grad_lambda || !grad_mu || !grad_rho))
    return fail(EFWI_ERR_ARG, "gradient outputs are NULL");
  Para P;
  int rc = read_para(para_fname, P);
  if (rc) return rc;
  JVal survey;
  if ((rc = read_json(P.survey, survey))) return rc;
  efwi_config cfg{};
  cfg.nz = P.nz; cfg.nx = P.nx; cfg.ny = P.ny; cfg.npml = P.npml; cfg.nt = P.nt;
  cfg.dz = P.dz; cfg.dx = P.dx; cfg.dy = P.dy; cfg.dt = P.dt; cfg.f0 = P.f0;
  cfg.formulation = 0;
  cfg.if_win = P.if_win ? 1 : 0;
  cfg.device = gpu_id;
  cfg.memory_mode = -1;  // the full ring when it fits, checkpointed segments otherwise
  std::vector<ShotDesc> shots(group_size);
  for (int k = 0; k < group_size; k++)
    if ((rc = read_shot(survey, shot_ids[k], P.npml, P.if_win, shots[k]))) return rc;

  efwi_ctx* raw = nullptr;
  if ((rc = efwi_create(&cfg, &raw))) return rc;
  std::unique_ptr<efwi_ctx, void (*)(efwi_ctx*)> ctx(raw, efwi_destroy);
  if ((rc = efwi_set_model(raw, lambda, mu, rho, 1, nullptr))) return rc;
  if ((rc = efwi_zero_gradient(raw, nullptr))) return rc;

  const int nt = P.nt;
  double total = 0.0;
  for (int k = 0; k < group_size; k++) {
    const ShotDesc& s = shots[k];
    const int nrec = (int)(s.rec.size() / 3);
    const size_t ns = (size_t)3 * nrec * nt;
    std::vector<float> host_obs, host_syn(ns), host_res;
    DevBuf d_obs, d_syn, d_res;
    if (hipMalloc(&d_syn.p, std::max<size_t>(ns, 1) * sizeof(float)) != hipSuccess)
      return fail(EFWI_ERR_NOMEM, "seismogram buffer");
    const std::string sid = std::to_string(s.id);
    if (calc_id != EFWI_CALC_SYNTHETIC) {
      host_obs.resize(ns);
      const char* comp[3] = {"x", "y", "z"};
      for (int c = 0; c < 3; c++)
        if ((rc = load_bin(P.data_dir + "/Shot_" + comp[c] + sid + ".bin", host_obs.data() + (size_t)c * nrec * nt,
                           (size_t)nrec * nt)))
          return rc;
      if (hipMalloc(&d_obs.p, std::max<size_t>(ns, 1) * sizeof(float)) != hipSuccess ||
          hipMalloc(&d_res.p, std::max<size_t>(ns, 1) * sizeof(float)) != hipSuccess)
        return fail(EFWI_ERR_NOMEM, "seismogram buffer");
      if (ns && hipMemcpy(d_obs.p, host_obs.data(), ns * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return fail(EFWI_ERR_HIP, "copy of observed data");
    }
    efwi_shot sh{};
    sh.src_z = s.sz; sh.src_x = s.sx; sh.src_y = s.sy;
    sh.stf = stf + (size_t)s.id * nt;  // stf row of the global shot id (Src/Src_Rec.cu:141)
    sh.nrec = nrec;
    sh.rec = s.rec.data();
    sh.win_start = s.ws.empty() ? nullptr : s.ws.data();
    sh.win_end = s.we.empty() ? nullptr : s.we.data();
    sh.weights = s.wt.empty() ? nullptr : s.wt.data();
    sh.src_weight = s.srcw;
    rc = efwi_run_shot(raw, &sh, calc_id, (const float*)d_obs.p, (float*)d_syn.p, (float*)d_res.p, &total,
                       nullptr);
    if (rc) return rc;
    if (hipDeviceSynchronize() != hipSuccess) return fail(EFWI_ERR_HIP, "device synchronize");
    const char* comp[3] = {"x", "y", "z"};
    if (ns && hipMemcpy(host_syn.data(), d_syn.p, ns * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
      return fail(EFWI_ERR_HIP, "copy of synthetic data");
    if (calc_id == EFWI_CALC_SYNTHETIC) {
      // Src/libCUFD.cu:704-725
      for (int c = 0; c < 3; c++)
        if ((rc = write_bin(P.data_dir + "/Shot_" + comp[c] + sid + ".bin", host_syn.data() + (size_t)c * nrec * nt,
                            (size_t)nrec * nt)))
          return rc;
    } else if (calc_id == EFWI_CALC_GRADIENT && P.save_scratch) {
      // Src/libCUFD.cu:660-701: windowed residual, windowed synthetic and
      // windowed observed data.
      host_res.resize(ns);
      if (ns && hipMemcpy(host_res.data(), d_res.p, ns * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(EFWI_ERR_HIP, "copy of residual");
      for (int c = 0; c < 3; c++)
        for (int r = 0; r < nrec; r++)
          for (int it = 0; it < nt; it++) {
            const size_t q = ((size_t)c * nrec + r) * nt + it;
            const float w = P.if_win ? taper_window_fac(it, nt, P.dt, 0.005f, s.ws[r], s.we[r],
                                                        s.wt.empty() ? 1.0f : s.wt[r], s.srcw)
                                     : taper_plain_w2(it, nt, P.dt, 0.005f);
            host_syn[q] = host_syn[q] * w;
            host_obs[q] = host_obs[q] * w;
          }
      for (int c = 0; c < 3; c++) {
        const size_t off = (size_t)c * nrec * nt, n = (size_t)nrec * nt;
        if ((rc = write_bin(P.scratch_dir + "/Residual_Shot_" + comp[c] + sid + ".bin", host_res.data() + off, n)) ||
            (rc = write_bin(P.scratch_dir + "/Syn_Shot_" + comp[c] + sid + ".bin", host_syn.data() + off, n)) ||
            (rc = write_bin(P.scratch_dir + "/CondObs_Shot_" + comp[c] + sid + ".bin", host_obs.data() + off, n)))
          return rc;
      }
    }
  }
  if (calc_id == EFWI_CALC_GRADIENT)
    if ((rc = efwi_get_gradient(raw, grad_lambda, grad_mu, grad_rho, 1, nullptr))) return rc;
  if (calc_id != EFWI_CALC_SYNTHETIC && misfit) *misfit = (float)total;
  return EFWI_OK;
}
