// .ori scene loader and writer.
//
// Reads the XML schema orion's Parser understands (util/parser.cpp:45-471)
// and resolves it into records: Model (meshHandy inline data or OBJ mesh)
// with an optional Transform chain, a Material (matte / metal / glass /
// mirror / plastic / uber, reusable through name= / named=), an optional
// AreaLight (one light per triangle), point / spot / distant Lights, Envmaps
// (8-bit PNG images or constant), the pinhole Camera, the Integrator and
// Sampler.  Values are parsed the way orion parses them (atof then a float
// cast, comma-separated lists, util/strutil.cpp:26-99), so the resolved
// floats equal the reference's.  The XML reader is a small one of this
// build's own (elements, attributes, comments, declarations).
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "host_math.h"
#include "host_scene.h"

namespace orion_amd {
namespace {

// ------------------------------------------------------------------ XML
struct XNode {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;  // document order
  std::vector<std::unique_ptr<XNode>> kids;
  const char *attr(const char *k) const {
    for (auto &a : attrs)
      if (a.first == k) return a.second.c_str();
    return nullptr;
  }
  const XNode *first(const char *n) const {
    for (auto &k : kids)
      if (k->name == n) return k.get();
    return nullptr;
  }
  std::vector<const XNode *> all(const char *n) const {
    std::vector<const XNode *> r;
    for (auto &k : kids)
      if (k->name == n) r.push_back(k.get());
    return r;
  }
};

struct XParser {
  const std::string &s;
  size_t i = 0;
  std::string err;
  explicit XParser(const std::string &src) : s(src) {}
  void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) ++i; }
  bool starts(const char *p) const { return s.compare(i, std::strlen(p), p) == 0; }
  bool skip_misc() {  // comments, declarations, processing instructions, text
    for (;;) {
      ws();
      if (starts("<!--")) {
        size_t e = s.find("-->", i);
        if (e == std::string::npos) { err = "unterminated comment"; return false; }
        i = e + 3;
      } else if (starts("<?")) {
        size_t e = s.find("?>", i);
        if (e == std::string::npos) { err = "unterminated declaration"; return false; }
        i = e + 2;
      } else if (starts("<!")) {
        size_t e = s.find('>', i);
        if (e == std::string::npos) { err = "unterminated markup"; return false; }
        i = e + 1;
      } else if (i < s.size() && s[i] != '<') {
        while (i < s.size() && s[i] != '<') ++i;  // character data (unused by the schema)
      } else {
        return true;
      }
    }
  }
  static std::string decode(const std::string &v) {
    std::string o;
    for (size_t k = 0; k < v.size(); ++k) {
      if (v[k] == '&') {
        size_t e = v.find(';', k);
        if (e != std::string::npos) {
          std::string ent = v.substr(k + 1, e - k - 1);
          if (ent == "amp") { o += '&'; k = e; continue; }
          if (ent == "lt") { o += '<'; k = e; continue; }
          if (ent == "gt") { o += '>'; k = e; continue; }
          if (ent == "quot") { o += '"'; k = e; continue; }
          if (ent == "apos") { o += '\''; k = e; continue; }
        }
      }
      o += v[k];
    }
    return o;
  }
  std::unique_ptr<XNode> element() {
    if (!skip_misc()) return nullptr;
    if (i >= s.size() || s[i] != '<') { err = "expected element"; return nullptr; }
    ++i;
    auto n = std::make_unique<XNode>();
    size_t b = i;
    while (i < s.size() && !std::isspace((unsigned char)s[i]) && s[i] != '>' && s[i] != '/') ++i;
    n->name = s.substr(b, i - b);
    for (;;) {
      ws();
      if (i >= s.size()) { err = "unterminated tag <" + n->name; return nullptr; }
      if (s[i] == '/') {
        if (i + 1 < s.size() && s[i + 1] == '>') { i += 2; return n; }
        err = "bad tag <" + n->name; return nullptr;
      }
      if (s[i] == '>') { ++i; break; }
      size_t kb = i;
      while (i < s.size() && !std::isspace((unsigned char)s[i]) && s[i] != '=') ++i;
      std::string key = s.substr(kb, i - kb);
      ws();
      if (i >= s.size() || s[i] != '=') { err = "expected = after " + key; return nullptr; }
      ++i;
      ws();
      if (i >= s.size() || (s[i] != '"' && s[i] != '\'')) { err = "expected quote"; return nullptr; }
      char q = s[i++];
      size_t vb = i;
      while (i < s.size() && s[i] != q) ++i;
      if (i >= s.size()) { err = "unterminated attribute"; return nullptr; }
      n->attrs.emplace_back(key, decode(s.substr(vb, i - vb)));
      ++i;
    }
    for (;;) {
      if (!skip_misc()) return nullptr;
      if (i >= s.size()) { err = "missing </" + n->name + ">"; return nullptr; }
      if (starts("</")) {
        i += 2;
        size_t b2 = i;
        while (i < s.size() && s[i] != '>') ++i;
        std::string close = s.substr(b2, i - b2);
        while (!close.empty() && std::isspace((unsigned char)close.back())) close.pop_back();
        if (close != n->name) { err = "mismatched </" + close + "> for <" + n->name + ">"; return nullptr; }
        ++i;
        return n;
      }
      auto k = element();
      if (!k) return nullptr;
      n->kids.push_back(std::move(k));
    }
  }
};

// ------------------------------------------------------------------ value parsing (strutil.cpp)
void trim(std::string &s) {  // strutil.cpp:4-10 (spaces only)
  if (s.empty()) return;
  s.erase(0, s.find_first_not_of(' '));
  s.erase(s.find_last_not_of(' ') + 1);
}
std::vector<std::string> split(const std::string &s) {  // strutil.cpp:11-25, delim ","
  size_t last = 0;
  size_t index = s.find_first_of(",", last);
  std::vector<std::string> res;
  while (index != std::string::npos) {
    res.push_back(s.substr(last, index - last));
    last = index + 1;
    index = s.find_first_of(",", last);
  }
  if (index - last > 0) res.push_back(s.substr(last, index - last));
  return res;
}
float pf(const std::string &s) { return static_cast<float>(std::atof(s.c_str())); }
int pi_(const std::string &s) { return std::atoi(s.c_str()); }

// ParamSet: first value of a key wins (std::map::insert, param.h:39-43)
struct ParamSet {
  std::map<std::string, std::string> kv;
  void set(const std::string &k, const std::string &v) { kv.insert(std::make_pair(k, v)); }
  bool has(const std::string &k) const { return kv.count(k) != 0; }
  std::string get(const std::string &k) const {
    auto it = kv.find(k);
    return it == kv.end() ? std::string() : it->second;
  }
};
ParamSet params_of(const XNode *n) {  // GET_PARAMSET, parser.h:20-29
  ParamSet ps;
  if (!n) return ps;
  for (auto *p : n->all("Param"))
    for (auto &a : p->attrs) ps.set(a.first, a.second);
  return ps;
}
std::vector<std::pair<std::string, std::string>> paramvec_of(const XNode *n) {  // GET_PARAMVEC
  std::vector<std::pair<std::string, std::string>> pv;
  if (!n) return pv;
  for (auto *p : n->all("Param"))
    for (auto &a : p->attrs) pv.push_back(a);
  return pv;
}

bool vec3(const std::string &s, float o[3]) {
  auto v = split(s);
  if (v.size() < 3) return false;
  for (int k = 0; k < 3; ++k) o[k] = pf(v[k]);
  return true;
}
void spectrum(const std::string &s, float o[3]) {  // strutil.cpp:77-84
  auto v = split(s);
  if (v.size() < 3) { o[0] = o[1] = o[2] = pf(v.empty() ? std::string() : v[0]); return; }
  for (int k = 0; k < 3; ++k) o[k] = pf(v[k]);
}

}  // namespace

// createTransform (transform.cpp:218-252): each entry left-multiplies.
bool create_transform(const std::vector<std::pair<std::string, std::string>> &pv, Xform *out,
                      std::string *err) {
  static Xform reuse[8];
  Xform t;
  for (auto &kv : pv) {
    const std::string &k = kv.first;
    if (k == "store") {
      int pos = pi_(kv.second);
      if (pos < 0 || pos >= 8) { *err = "store index out of range"; return false; }
      reuse[pos] = t;
    } else if (k == "reuse") {
      int pos = pi_(kv.second);
      if (pos < 0 || pos >= 8) { *err = "reuse index out of range"; return false; }
      t = reuse[pos] * t;
    } else if (k == "scale") {
      float v[3];
      if (!vec3(kv.second, v)) { *err = "scale needs 3 values"; return false; }
      t = scale(v[0], v[1], v[2]) * t;
    } else if (k == "rotate") {
      auto v = split(kv.second);
      if (v.size() <= 3) { *err = "rotate needs 4 values"; return false; }
      t = rotate(pf(v[3]), f3(pf(v[0]), pf(v[1]), pf(v[2]))) * t;
    } else if (k == "translate") {
      float v[3];
      if (!vec3(kv.second, v)) { *err = "translate needs 3 values"; return false; }
      t = translate(f3(v[0], v[1], v[2])) * t;
    } else if (k == "lookat") {
      auto v = split(kv.second);
      if (v.size() < 9) { *err = "lookat needs 9 values"; return false; }
      t = look_at(f3(pf(v[0]), pf(v[1]), pf(v[2])), f3(pf(v[3]), pf(v[4]), pf(v[5])),
                  f3(pf(v[6]), pf(v[7]), pf(v[8]))) * t;
    }
  }
  *out = t;
  return true;
}

namespace {

struct Mesh {  // de-indexed corners, core/triangle.h:16-25
  std::vector<float> p, n;  // 3 per corner
};

// meshHandy (strutil.cpp:101-159): p required, n either one normal or per vertex.
bool mesh_handy(const ParamSet &ps, Mesh *m, std::string *err) {
  auto ps_ = split(ps.get("p"));
  if (ps_.size() % 3 != 0) { *err = "meshHandy: p is not a list of points"; return false; }
  int nv = (int)ps_.size() / 3;
  if (nv % 3 != 0) { *err = "meshHandy: vertex count is not a multiple of 3"; return false; }
  m->p.resize(3 * nv);
  for (int i = 0; i < 3 * nv; ++i) m->p[i] = pf(ps_[i]);
  if (!ps.has("n")) { *err = "meshHandy without n: orion reads normals unconditionally"; return false; }
  auto ns = split(ps.get("n"));
  m->n.resize(3 * nv);
  if (ns.size() == 3) {
    for (int i = 0; i < nv; ++i)
      for (int k = 0; k < 3; ++k) m->n[3 * i + k] = pf(ns[k]);
  } else {
    if ((int)ns.size() < 3 * nv) { *err = "meshHandy: too few normals"; return false; }
    for (int i = 0; i < 3 * nv; ++i) m->n[i] = pf(ns[i]);
  }
  return true;
}

// OBJ (triangle.cpp:8-58 through tiny_obj_loader): v / vn / f with v//vn or
// v/vt/vn corners; every corner needs a normal; polygons must be triangles.
bool load_obj(const std::string &path, Mesh *m, std::string *err) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) { *err = "cannot open " + path; return false; }
  std::vector<float> v, vn;
  char line[4096];
  auto idx = [](long k, size_t n) -> long { return k > 0 ? k - 1 : (long)n + k; };
  bool ok = true;
  while (std::fgets(line, sizeof(line), f)) {
    char *s = line;
    while (*s == ' ' || *s == '\t') ++s;
    if (s[0] == 'v' && (s[1] == ' ' || s[1] == '\t')) {
      float a, b, c;
      if (std::sscanf(s + 2, "%f %f %f", &a, &b, &c) != 3) { ok = false; *err = "bad v line"; break; }
      v.push_back(a); v.push_back(b); v.push_back(c);
    } else if (s[0] == 'v' && s[1] == 'n') {
      float a, b, c;
      if (std::sscanf(s + 3, "%f %f %f", &a, &b, &c) != 3) { ok = false; *err = "bad vn line"; break; }
      vn.push_back(a); vn.push_back(b); vn.push_back(c);
    } else if (s[0] == 'f' && (s[1] == ' ' || s[1] == '\t')) {
      long vi[3], ni[3];
      char *q = s + 2;
      int c = 0;
      while (*q && c < 4) {
        while (*q == ' ' || *q == '\t') ++q;
        if (!*q || *q == '\n' || *q == '\r') break;
        char *e;
        long a = std::strtol(q, &e, 10);
        long nn = 0;
        if (*e == '/') {
          ++e;
          if (*e != '/') std::strtol(e, &e, 10);  // texcoord (unused)
          if (*e == '/') { ++e; nn = std::strtol(e, &e, 10); }
        }
        if (c < 3) { vi[c] = a; ni[c] = nn; }
        ++c;
        q = e;
        while (*q && *q != ' ' && *q != '\t' && *q != '\n') ++q;
      }
      if (c != 3) { ok = false; *err = "only triangles are supported"; break; }
      for (int k = 0; k < 3; ++k) {
        if (ni[k] == 0) { ok = false; *err = "OBJ corner without normal"; break; }
        long a = idx(vi[k], v.size() / 3), b = idx(ni[k], vn.size() / 3);
        if (a < 0 || (size_t)a >= v.size() / 3 || b < 0 || (size_t)b >= vn.size() / 3) {
          ok = false; *err = "OBJ index out of range"; break;
        }
        for (int d = 0; d < 3; ++d) { m->p.push_back(v[3 * a + d]); m->n.push_back(vn[3 * b + d]); }
      }
      if (!ok) break;
    }
  }
  std::fclose(f);
  return ok;
}

struct MatSpec { std::vector<float> rec; };

bool texture_color(const XNode *mat, const char *name, bool *present, float o[3]) {
  const XNode *n = mat->first(name);
  *present = false;
  if (!n) return true;
  ParamSet ps = params_of(n);
  std::string type = ps.get("type");
  trim(type);
  if (type == "constant" && !ps.get("color").empty()) {
    spectrum(ps.get("color"), o);
    *present = true;
  } else if (type == "float" && !ps.get("value").empty()) {
    float v = pf(ps.get("value"));
    o[0] = o[1] = o[2] = v;
    *present = true;
  }
  return true;
}

// A FloatTexture parameter (CREATE_FLOAT_TEXTURE, parser.cpp:37-43): only
// type "float" survives the dynamic_pointer_cast; anything else is absent.
bool float_texture(const XNode *mat, const char *name, bool *present, float *v) {
  const XNode *n = mat->first(name);
  *present = false;
  if (!n) return true;
  ParamSet ps = params_of(n);
  std::string type = ps.get("type");
  trim(type);
  if (type == "float" && !ps.get("value").empty()) {
    *v = pf(ps.get("value"));
    *present = true;
  }
  return true;
}

// Parser::_makeMaterial (parser.cpp:262-431)
bool make_material(const XNode *mat, std::map<std::string, std::vector<float>> &named,
                   std::vector<float> *rec, std::string *err) {
  if (!mat) { *err = "Model without Material"; return false; }
  const char *nm = mat->attr("named");
  rec->assign(kMatFloats, 0.f);
  if (nm && *nm) {
    auto it = named.find(nm);
    if (it == named.end()) { *err = std::string("no named material ") + nm; return false; }
    *rec = it->second;
  } else {
    std::string type = mat->attr("type") ? mat->attr("type") : "";
    bool have;
    if (type == "matte") {
      float kd[3] = {0, 0, 0}, sg[3] = {0, 0, 0};
      texture_color(mat, "Kd", &have, kd);
      if (!have) { *err = "matte without Kd"; return false; }
      // matte.cpp:23-29: a missing (or non-float) sigma is FloatTexture(0)
      float_texture(mat, "sigma", &have, sg);
      matte_record(kd[0], kd[1], kd[2], rec->data(), have ? sg[0] : 0.f);
    } else if (type == "metal") {
      float eta[3] = {0, 0, 0}, k[3] = {0, 0, 0}, r[3] = {0.1f, 0.1f, 0.1f};
      bool he, hk, hr;
      texture_color(mat, "eta", &he, eta);
      texture_color(mat, "k", &hk, k);
      texture_color(mat, "roughness", &hr, r);
      // metal.cpp:74-140: explicit eta/k only when all three textures exist
      if (!(he && hk && hr)) {
        ParamSet ps = params_of(mat);
        if (!ps.has("metalName") || !metal_preset(ps.get("metalName"), eta, k)) {
          *err = "metal needs eta, k and roughness or a known metalName"; return false;
        }
        if (!hr) r[0] = 0.1f;
      }
      metal_record(eta, k, r[0], rec->data());
    } else if (type == "glass") {
      float kr[3] = {0, 0, 0}, kt[3] = {0, 0, 0}, eta[3] = {1.5f, 1.5f, 1.5f}, r[3] = {0, 0, 0};
      bool h;
      texture_color(mat, "Kr", &h, kr);
      texture_color(mat, "Kt", &h, kt);
      texture_color(mat, "eta", &h, eta);
      if (!h) { *err = "glass without eta"; return false; }
      texture_color(mat, "roughness", &h, r);
      if (!h) { *err = "glass without roughness"; return false; }
      glass_record(kr, kt, r[0], eta[0], rec->data());
    } else if (type == "mirror") {  // parser.cpp:297-307
      float kr[3] = {0, 0, 0};
      texture_color(mat, "Kr", &have, kr);
      if (!have) { *err = "mirror without Kr"; return false; }
      mirror_record(kr, rec->data());
    } else if (type == "plastic") {  // parser.cpp:339-365
      float kd[3] = {0, 0, 0}, ks[3] = {0, 0, 0}, r = 0;
      texture_color(mat, "Kd", &have, kd);
      if (!have) { *err = "plastic without Kd"; return false; }
      texture_color(mat, "Ks", &have, ks);
      if (!have) { *err = "plastic without Ks"; return false; }
      float_texture(mat, "roughness", &have, &r);
      if (!have) { *err = "plastic without a float roughness"; return false; }
      plastic_record(kd, ks, r, rec->data());
    } else if (type == "uber") {  // parser.cpp:412-421: every texture is evaluated
      float kd[3], ks[3], kr[3], kt[3], op[3], r = 0, e = 0;
      const char *names[5] = {"Kd", "Ks", "Kr", "Kt", "opacity"};
      float *dst[5] = {kd, ks, kr, kt, op};
      for (int i = 0; i < 5; ++i) {
        dst[i][0] = dst[i][1] = dst[i][2] = 0;
        texture_color(mat, names[i], &have, dst[i]);
        if (!have) { *err = std::string("uber without ") + names[i]; return false; }
      }
      float_texture(mat, "roughness", &have, &r);
      if (!have) { *err = "uber without a float roughness"; return false; }
      float_texture(mat, "eta", &have, &e);
      if (!have) { *err = "uber without a float eta"; return false; }
      uber_record(kd, ks, kr, kt, op, r, e, rec->data());
    } else {
      *err = "material type '" + type + "' is outside this build's scope";
      return false;
    }
  }
  const char *name = mat->attr("name");
  if (name && *name) named.insert(std::make_pair(std::string(name), *rec));
  return true;
}

std::string dirname_of(const std::string &p) {
  size_t k = p.find_last_of('/');
  return k == std::string::npos ? std::string() : p.substr(0, k + 1);
}

}  // namespace

int load_ori(const std::string &path, HostScene *out, std::string *err) {
  std::ifstream in(path, std::ios::binary);
  if (!in) { *err = "cannot open " + path; return ORION_ERR_IO; }
  std::stringstream ss;
  ss << in.rdbuf();
  std::string src = ss.str();
  XParser xp(src);
  auto root = xp.element();
  if (!root) { *err = "XML: " + xp.err; return ORION_ERR_PARSE; }
  HostScene s;
  // Resource directory: relative to the .ori's folder (orion resolves it
  // against the executable's folder, envvariable.cpp:37-40).
  std::string res = dirname_of(path);
  if (const XNode *r = root->first("Resource")) {
    const char *fn = r->attr("filename");
    if (fn) {
      std::string f = fn;
      trim(f);
      res = (!f.empty() && f[0] == '/') ? f : dirname_of(path) + f;
      if (!res.empty() && res.back() != '/') res += '/';
    }
  }
  // Camera (parser.cpp:139-172)
  const XNode *cam = root->first("Camera");
  if (!cam) { *err = "No camera!"; return ORION_ERR_PARSE; }
  {
    std::string type = cam->attr("type") ? cam->attr("type") : "";
    trim(type);
    if (type != "pinhole") { *err = "camera type '" + type + "' unsupported"; return ORION_ERR_PARSE; }
    ParamSet cp = params_of(cam);
    Xform t;
    if (!create_transform(paramvec_of(cam->first("Transform")), &t, err)) return ORION_ERR_PARSE;
    auto fs = split(cp.get("filmsize"));
    if (fs.size() < 2) { *err = "camera needs filmsize"; return ORION_ERR_PARSE; }
    int W = pi_(fs[0]), H = pi_(fs[1]);
    if (W <= 0 || H <= 0) { *err = "bad filmsize"; return ORION_ERR_PARSE; }
    float fov = pf(cp.get("fov"));
    if (cp.has("lensRadius") && pf(cp.get("lensRadius")) > 0) {
      *err = "thin-lens cameras are outside this build's scope"; return ORION_ERR_UNSUPPORTED;
    }
    F3 o = apply_point(t.m, f3(0, 0, 0));
    s.camera[0] = o.x; s.camera[1] = o.y; s.camera[2] = o.z;
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) s.camera[3 + 3 * r + c] = t.m.m[r][c];
    s.camera[12] = std::tan(radians(fov * 0.5f));
    s.camera[13] = (float)W / H;
    s.camera[14] = (float)W;
    s.camera[15] = (float)H;
    s.fov = fov;
    auto la = split([&] {
      for (auto &kv : paramvec_of(cam->first("Transform")))
        if (kv.first == "lookat") return kv.second;
      return std::string();
    }());
    if (la.size() >= 9)
      for (int k = 0; k < 9; ++k) s.lookat[k] = pf(la[k]);
  }
  // Models (parser.cpp:58-121): area lights are pushed before <Light> lights.
  std::map<std::string, std::vector<float>> named;
  for (const XNode *model : root->all("Model")) {
    Xform t;
    if (!create_transform(paramvec_of(model->first("Transform")), &t, err)) return ORION_ERR_PARSE;
    std::string type = model->attr("type") ? model->attr("type") : "";
    Mesh mesh;
    if (type == "mesh") {
      const char *fn = model->attr("filename");
      if (!fn) { *err = "mesh without filename"; return ORION_ERR_PARSE; }
      if (!load_obj(res + fn, &mesh, err)) return ORION_ERR_IO;
    } else if (type == "meshHandy") {
      if (!mesh_handy(params_of(model), &mesh, err)) return ORION_ERR_PARSE;
    } else {
      *err = "model type '" + type + "' unsupported";
      return ORION_ERR_PARSE;
    }
    // createTriangleMesh (triangle.cpp:137-151): world-space p and n
    size_t nc = mesh.p.size() / 3;
    for (size_t c = 0; c < nc; ++c) {
      F3 n = apply_normal(t, f3(mesh.n[3 * c], mesh.n[3 * c + 1], mesh.n[3 * c + 2]));
      F3 p = apply_point(t.m, f3(mesh.p[3 * c], mesh.p[3 * c + 1], mesh.p[3 * c + 2]));
      mesh.p[3 * c] = p.x; mesh.p[3 * c + 1] = p.y; mesh.p[3 * c + 2] = p.z;
      mesh.n[3 * c] = n.x; mesh.n[3 * c + 1] = n.y; mesh.n[3 * c + 2] = n.z;
    }
    std::vector<float> mrec;
    if (!make_material(model->first("Material"), named, &mrec, err)) return ORION_ERR_PARSE;
    int mat_index = s.n_mats();
    s.mats.insert(s.mats.end(), mrec.begin(), mrec.end());
    const XNode *al = model->first("AreaLight");
    float Le[3] = {0, 0, 0};
    if (al) spectrum(params_of(al).get("Le"), Le);
    int ntri = (int)(nc / 3);
    for (int k = 0; k < ntri; ++k) {
      int tri = s.n_tris();
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) s.tri_pn.push_back(mesh.p[9 * k + 3 * c + d]);
      for (int c = 0; c < 3; ++c)
        for (int d = 0; d < 3; ++d) s.tri_pn.push_back(mesh.n[9 * k + 3 * c + d]);
      s.tri_mat.push_back(mat_index);
      if (al) {
        if (tri >= (1 << 24)) { *err = "area light on triangle index >= 2^24"; return ORION_ERR_UNSUPPORTED; }
        s.tri_light.push_back(s.n_lights());
        float rec[32];
        area_light_record(tri, Le, rec);
        s.lights.insert(s.lights.end(), rec, rec + 32);
      } else {
        s.tri_light.push_back(-1);
      }
    }
  }
  // Lights (parser.cpp:174-194)
  for (const XNode *ln : root->all("Light")) {
    std::string type = ln->attr("type") ? ln->attr("type") : "";
    trim(type);
    ParamSet ps = params_of(ln);
    float rec[32];
    if (type == "point") {
      float I[3], from[3];
      spectrum(ps.get("I"), I);
      if (!vec3(ps.get("from"), from)) { *err = "point light needs from"; return ORION_ERR_PARSE; }
      point_light_record(I, from, rec);
    } else if (type == "spot") {
      float I[3], from[3], to[3];
      spectrum(ps.get("intensity"), I);
      if (!vec3(ps.get("from"), from) || !vec3(ps.get("to"), to)) {
        *err = "spot light needs from and to"; return ORION_ERR_PARSE;
      }
      spot_light_record(I, from, to, pf(ps.get("angle")), pf(ps.get("delta")), rec);
    } else if (type == "distant") {  // distant.cpp:58-65
      float L[3], from[3], to[3];
      spectrum(ps.get("L"), L);
      if (!vec3(ps.get("from"), from) || !vec3(ps.get("to"), to)) {
        *err = "distant light needs from and to"; return ORION_ERR_PARSE;
      }
      distant_light_record(L, from, to, rec);
    } else {
      *err = "light type '" + type + "' is outside this build's scope";
      return ORION_ERR_UNSUPPORTED;
    }
    s.lights.insert(s.lights.end(), rec, rec + 32);
  }
  // Envmaps (parser.cpp:221-242, createEnvMap envmap.cpp:108-128), after
  // every other light
  for (const XNode *en : root->all("Envmap")) {
    EnvImage img;
    img.xform = paramvec_of(en->first("Transform"));
    Xform t;
    if (!create_transform(img.xform, &t, err)) return ORION_ERR_PARSE;
    ParamSet ps = params_of(en);
    float L[3] = {1.f, 1.f, 1.f}, up[3];
    if (ps.has("L")) spectrum(ps.get("L"), L);
    // parseVector3f on a missing worldUp reads past an empty vector in the
    // reference: required here
    if (!vec3(ps.get("worldUp"), up)) { *err = "Envmap needs worldUp"; return ORION_ERR_PARSE; }
    if (ps.has("name")) {
      img.file = ps.get("name");
      int w, h, ch;
      std::vector<unsigned char> px;
      if (!png_read(res + img.file, &w, &h, &ch, &px, err)) return ORION_ERR_IO;
      if (ch != 3 && ch != 4) { *err = img.file + ": only 3 or 4 channel images are supported"; return ORION_ERR_UNSUPPORTED; }
      img.width = w;
      img.height = h;
      img.rgb.resize((size_t)w * h * 3);
      const float inv255 = 0.0039215686274509803921568627451f;  // imagetexture.cpp:15-20
      for (size_t i = 0; i < (size_t)w * h; ++i)
        for (int c = 0; c < 3; ++c) img.rgb[3 * i + c] = px[ch * i + c] * inv255;
    } else {
      img.rgb.assign(3, 1.f);  // ConstantTexture(L): a 1 x 1 map of L
    }
    float l2w[16], w2l[16];
    envmap_orientation(&t.m.m[0][0], &t.mInv.m[0][0], up, l2w, w2l);
    float rec[32];
    envmap_light_record((int)s.env.size(), L, img.width, img.height, pi_(ps.get("nLightSamples")), l2w, w2l, up,
                        rec);
    s.lights.insert(s.lights.end(), rec, rec + 32);
    s.env.push_back(std::move(img));
  }
  // Sampler (parser.cpp:196-212)
  if (const XNode *sm = root->first("Sampler")) {
    std::string type = sm->attr("type") ? sm->attr("type") : "";
    if (type != "sobol") { *err = "sampler '" + type + "' unsupported (sobol only)"; return ORION_ERR_UNSUPPORTED; }
  }
  // Integrator (parser.cpp:123-137, path.cpp:108-113, sppm.cpp:324-333)
  const XNode *ig = root->first("Integrator");
  if (!ig) { *err = "No integrator"; return ORION_ERR_PARSE; }
  {
    std::string type = ig->attr("type") ? ig->attr("type") : "";
    trim(type);
    ParamSet ps = params_of(ig);
    s.info.max_depth = pi_(ps.get("maxDepth"));
    if (type == "path") {
      s.info.integrator = 0;
      s.info.spp = pi_(ps.get("nSamplesPerPixel"));
    } else if (type == "sppm") {
      s.info.integrator = 1;
      s.info.sppm_iterations = pi_(ps.get("numIterations"));
      s.info.sppm_photons = pi_(ps.get("photonsPerIteration"));
      s.info.sppm_radius = pf(ps.get("initRadius"));
    } else {
      *err = "integrator '" + type + "' unsupported";
      return ORION_ERR_UNSUPPORTED;
    }
  }
  s.refresh_info();
  *out = std::move(s);
  return ORION_OK;
}

// ------------------------------------------------------------------ writer
namespace {
std::string fstr(float v) {
  char b[32];
  std::snprintf(b, sizeof(b), "%.9g", (double)v);
  return b;
}
std::string f3str(const float *v) { return fstr(v[0]) + "," + fstr(v[1]) + "," + fstr(v[2]); }
void write_material(std::ostream &o, const float *m) {
  int type = (int)m[0];
  auto cst = [&](const char *n, const float *v) {
    o << "      <" << n << "><Param type=\"constant\"/><Param color=\"" << f3str(v) << "\"/></" << n << ">\n";
  };
  auto flt = [&](const char *n, float v) {
    o << "      <" << n << "><Param type=\"float\"/><Param value=\"" << fstr(v) << "\"/></" << n << ">\n";
  };
  if (type == MAT_MATTE) {
    o << "    <Material type=\"matte\">\n";
    cst("Kd", m + 1);
    if (m[4] != 0) flt("sigma", m[4]);
  } else if (type == MAT_MIRROR) {
    o << "    <Material type=\"mirror\">\n";
    cst("Kr", m + 1);
  } else if (type == MAT_PLASTIC) {
    o << "    <Material type=\"plastic\">\n";
    cst("Kd", m + 1);
    cst("Ks", m + 4);
    flt("roughness", m[7]);
  } else if (type == MAT_UBER) {
    o << "    <Material type=\"uber\">\n";
    cst("Kd", m + 1);
    cst("Ks", m + 4);
    cst("Kr", m + 7);
    cst("Kt", m + 10);
    cst("opacity", m + 13);
    flt("roughness", m[16]);
    flt("eta", m[18]);
  } else if (type == MAT_METAL) {
    o << "    <Material type=\"metal\">\n";
    cst("eta", m + 1);
    cst("k", m + 4);
    flt("roughness", m[7]);
  } else {
    o << "    <Material type=\"glass\">\n";
    cst("Kr", m + 1);
    cst("Kt", m + 4);
    flt("eta", m[9]);
    flt("roughness", m[7]);
  }
  o << "    </Material>\n";
}
}  // namespace

// Writes one Model per (material, area-light emission) run of consecutive
// triangles, so loading the file back reproduces the same triangle order,
// materials and light order.
int write_ori(const HostScene &s, const std::string &path, int inline_limit, std::string *err) {
  std::ofstream o(path);
  if (!o) { *err = "cannot write " + path; return ORION_ERR_IO; }
  std::string base = path;
  size_t dot = base.find_last_of('.');
  if (dot != std::string::npos) base = base.substr(0, dot);
  std::string stem = base.substr(base.find_last_of('/') + 1);
  o << "<Scene>\n  <Resource filename=\"./\"/>\n";
  o << "  <Camera type=\"pinhole\">\n    <Transform><Param lookat=\"";
  for (int k = 0; k < 9; ++k) o << (k ? "," : "") << fstr(s.lookat[k]);
  o << "\"/></Transform>\n    <Param fov=\"" << fstr(s.fov) << "\"/>\n    <Param filmsize=\""
    << (int)s.camera[14] << "," << (int)s.camera[15] << "\"/>\n  </Camera>\n";
  int n = s.n_tris();
  int run = 0;
  for (int a = 0; a < n;) {
    int m = s.tri_mat[a];
    bool lit = s.tri_light[a] >= 0;
    const float *Le = lit ? &s.lights[32 * s.tri_light[a] + 2] : nullptr;
    int b = a + 1;
    while (b < n && s.tri_mat[b] == m && (s.tri_light[b] >= 0) == lit &&
           (!lit || std::memcmp(&s.lights[32 * s.tri_light[b] + 2], Le, 12) == 0))
      ++b;
    int cnt = b - a;
    if (cnt > inline_limit) {
      std::string objname = stem + "_" + std::to_string(run) + ".obj";
      std::string objpath = base.substr(0, base.find_last_of('/') + 1) + objname;
      FILE *f = std::fopen(objpath.c_str(), "w");
      if (!f) { *err = "cannot write " + objpath; return ORION_ERR_IO; }
      for (int t = a; t < b; ++t)
        for (int c = 0; c < 3; ++c) {
          const float *p = &s.tri_pn[18 * (size_t)t + 3 * c];
          std::fprintf(f, "v %.9g %.9g %.9g\n", p[0], p[1], p[2]);
        }
      for (int t = a; t < b; ++t)
        for (int c = 0; c < 3; ++c) {
          const float *q = &s.tri_pn[18 * (size_t)t + 9 + 3 * c];
          std::fprintf(f, "vn %.9g %.9g %.9g\n", q[0], q[1], q[2]);
        }
      for (int t = 0; t < cnt; ++t) {
        long v0 = 3L * t + 1;
        std::fprintf(f, "f %ld//%ld %ld//%ld %ld//%ld\n", v0, v0, v0 + 1, v0 + 1, v0 + 2, v0 + 2);
      }
      std::fclose(f);
      o << "  <Model type=\"mesh\" filename=\"" << objname << "\">\n";
    } else {
      o << "  <Model type=\"meshHandy\">\n    <Param p=\"";
      for (int t = a; t < b; ++t)
        for (int c = 0; c < 9; ++c) o << ((t == a && c == 0) ? "" : ",") << fstr(s.tri_pn[18 * (size_t)t + c]);
      o << "\"/>\n    <Param n=\"";
      for (int t = a; t < b; ++t)
        for (int c = 0; c < 9; ++c) o << ((t == a && c == 0) ? "" : ",") << fstr(s.tri_pn[18 * (size_t)t + 9 + c]);
      o << "\"/>\n";
    }
    write_material(o, &s.mats[kMatFloats * (size_t)m]);
    if (lit) o << "    <AreaLight><Param Le=\"" << f3str(Le) << "\"/><Param nLightSamples=\"1\"/></AreaLight>\n";
    o << "  </Model>\n";
    a = b;
    ++run;
  }
  for (int l = 0; l < s.n_lights(); ++l) {
    const float *r = &s.lights[32 * (size_t)l];
    if ((int)r[0] == LIGHT_POINT) {
      o << "  <Light type=\"point\"><Param I=\"" << f3str(r + 2) << "\"/><Param from=\"" << f3str(r + 5)
        << "\"/></Light>\n";
    } else if ((int)r[0] == LIGHT_SPOT) {
      const float *to = r + 29;  // loader inputs kept in the record
      float ang = r[8];
      o << "  <Light type=\"spot\"><Param intensity=\"" << f3str(r + 2) << "\"/><Param from=\""
        << f3str(r + 5) << "\"/><Param to=\"" << f3str(to) << "\"/><Param angle=\"" << fstr(ang)
        << "\"/><Param delta=\"0\"/></Light>\n";
    } else if ((int)r[0] == LIGHT_DISTANT) {
      o << "  <Light type=\"distant\"><Param L=\"" << f3str(r + 2) << "\"/><Param from=\"" << f3str(r + 8)
        << "\"/><Param to=\"" << f3str(r + 11) << "\"/></Light>\n";
    }
  }
  for (int l = 0; l < s.n_lights(); ++l) {
    const float *r = &s.lights[32 * (size_t)l];
    if ((int)r[0] != LIGHT_ENVMAP) continue;
    const EnvImage &img = s.env[(size_t)r[1]];
    if (!img.has_xform_params) { *err = "envmap orientation given as matrices cannot be written as .ori"; return ORION_ERR_UNSUPPORTED; }
    o << "  <Envmap>\n";
    if (!img.xform.empty()) {
      o << "    <Transform>";
      for (auto &kv : img.xform) o << "<Param " << kv.first << "=\"" << kv.second << "\"/>";
      o << "</Transform>\n";
    }
    o << "    <Param L=\"" << f3str(r + 2) << "\"/>\n";
    bool constant = img.file.empty() && img.width * img.height == 1 && img.rgb[0] == 1.f && img.rgb[1] == 1.f &&
                    img.rgb[2] == 1.f;
    if (!constant) {
      // texels are byte / 255 values: written back as an 8-bit PNG beside the .ori
      std::string fn = img.file.empty() ? stem + "_env" + std::to_string(l) + ".png" : img.file;
      size_t slash = fn.find_last_of('/');
      std::string flat = slash == std::string::npos ? fn : fn.substr(slash + 1);
      std::vector<unsigned char> px(img.rgb.size());
      for (size_t i = 0; i < px.size(); ++i) {
        float v = std::round(img.rgb[i] * 255.f);
        px[i] = (unsigned char)(v < 0 ? 0 : (v > 255 ? 255 : v));
      }
      std::string dir = base.substr(0, base.size() - stem.size());
      if (!png_write_rgb8(dir + flat, img.width, img.height, px.data(), err)) return ORION_ERR_IO;
      o << "    <Param name=\"" << flat << "\"/>\n";
    }
    o << "    <Param nLightSamples=\"" << (int)r[7] << "\"/>\n";
    o << "    <Param worldUp=\"" << f3str(r + 26) << "\"/>\n";
    o << "  </Envmap>\n";
  }
  const orion_scene_info &in = s.info;
  if (in.integrator == 1) {
    o << "  <Integrator type=\"sppm\"><Param maxDepth=\"" << in.max_depth << "\"/><Param numIterations=\""
      << in.sppm_iterations << "\"/><Param photonsPerIteration=\"" << in.sppm_photons
      << "\"/><Param initRadius=\"" << fstr(in.sppm_radius) << "\"/></Integrator>\n";
  } else {
    o << "  <Integrator type=\"path\"><Param maxDepth=\"" << in.max_depth << "\"/><Param nSamplesPerPixel=\""
      << in.spp << "\"/></Integrator>\n";
  }
  o << "  <Sampler type=\"sobol\"/>\n</Scene>\n";
  return o ? ORION_OK : ORION_ERR_IO;
}

}  // namespace orion_amd
